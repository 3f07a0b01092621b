! Drop-in replacement of the reference executable built from src/omp_parallel/ibm_2d_backstep_omp_cpu.f90:
! run it in a case directory (config/controlDict.txt, the porosity CSV); the time step runs
! on the GPU through libpixelflow.
program main
  use pf_driver
  implicit none
  call run_case('2d_backstep')
end program main
