! Drop-in replacement of the reference executable built from src/omp_parallel/ibm_3d_uniform_omp_cpu.f90:
! run it in a case directory (config/controlDict.txt, the porosity CSV); the time step runs
! on the GPU through libpixelflow.
program main
  use pf_driver
  implicit none
  call run_case('3d_uniform')
end program main
