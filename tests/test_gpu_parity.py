"""GPU parity: the HIP path against the reference's own outputs and the oracle.

* golden fixtures written by the reference build: bit for bit (u, v, w, p, the SOR error
  diagnostic of every step, the divergence diagnostic);
* the oracle on seeded inputs at larger sizes, odd extents and the air-conditioning
  solver: bit for bit;
* the 3D benchmark grid (256x128x128 sphere) for two full steps: rel-L2 <= 1e-12 (the
  stated bar; the result is in fact bit-identical).
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden
from pixelflow_amd import cases

pytestmark = pytest.mark.gpu

# the full-length benchmark fixtures have their own tests (tests/test_gpu_longrun.py)
FULL = [n for n in golden_names() if not n.startswith("g3d_ref_") and n not in ("g2d_c1_2000", "g3d_c2_500")]


def _solver(p, phi):
    from pixelflow_amd.solver import Solver
    return Solver(p, phi)


def _fields(p):
    return ("u", "v", "w", "p") if p.dim == 3 else ("u", "v", "p")


@pytest.mark.parametrize("name", FULL)
def test_gpu_matches_reference_fixture(name):
    p, g = load_golden(name)
    print(f"[gpu] fixture {name} grid={p.m}x{p.n}x{p.l} steps={p.istep_max} iter={p.iter_max}", flush=True)
    with _solver(p, g["phi_raw"]) as s:
        st = s.step(p.istep_max)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), g[f], err_msg=f"{name}:{f}")
        np.testing.assert_array_equal(st["sor_error"], g["errs"], err_msg=f"{name}:errs")
        np.testing.assert_array_equal(s.divergence(), g["div"], err_msg=f"{name}:div")


def test_gpu_matches_reference_room_project():
    import hashlib
    p, g = load_golden("g3d_ref_room64")
    with _solver(p, g["phi_raw"]) as s:
        st = s.step(p.istep_max)
        np.testing.assert_array_equal(st["sor_error"], g["errs"])
        for f in ("u", "v", "w", "p"):
            a = s.interior(f)
            assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest() == str(g["sha_" + f]), f


CASES = {
    "2d_c1_256x128": lambda: cases.cylinder_case(steps=8),
    "2d_odd_97x51": lambda: (cases.CaseParams(dim=2, m=97, n=51, width=1.0, height=0.52, time=6e-3, istep_max=6,
                                              iter_max=40, xlambda=1e-4),
                             cases.cylinder_2d(97, 51, 1.0, 0.52, (0.3, 0.26), 0.15)),
    "3d_sphere_64x32x32": lambda: cases.sphere_case(m=64, n=32, l=32, steps=4),
    "3d_odd_45x27x19": lambda: (cases.CaseParams(dim=3, m=45, n=27, l=19, width=1.0, height=0.59, depth=0.41,
                                                 time=3e-3, istep_max=3, iter_max=30),
                                cases.sphere_3d(45, 27, 19, 1.0, 0.59, 0.41, (0.3, 0.3, 0.2), 0.2)),
    "3d_boxes_48x32x24": lambda: cases.boxes_case(m=48, n=32, l=24, steps=3, iter_max=50, nboxes=4, side_range=(3, 10)),
    "3d_air_room_40x36x32": lambda: cases.room_case(4, steps=4, iter_max=50, size=(40, 36, 32)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_oracle(name):
    import pyoracle
    p, phi = CASES[name]()
    print(f"[gpu] oracle case {name} steps={p.istep_max}", flush=True)
    o = pyoracle.OracleRun(p, phi)
    e_ref = o.step(p.istep_max)
    with _solver(p, phi) as s:
        st = s.step(p.istep_max)
        np.testing.assert_array_equal(st["sor_error"], e_ref)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), o.interior(f), err_msg=f"{name}:{f}")
            # ghosts too: the state the next step reads
            full = s.field(f)
            ref = {"u": o.u, "v": o.v, "w": o.w, "p": o.p_}[f]
            if p.dim == 3:
                np.testing.assert_array_equal(full[1:-1, 1:-1, :], ref[1:-1, 1:-1, :], err_msg=f"{name}:{f} x-ghosts")
            else:
                np.testing.assert_array_equal(full[1:-1, :], ref[1:-1, :], err_msg=f"{name}:{f} x-ghosts")



def _extreme_velocities(p, rng):
    """Velocity fields whose magnitudes span 1e-305 .. 1e120 (random signs, zeros of both
    signs), with a band of ordinary values: the predictor's dividends then fall below and
    above the range where v_div_scale leaves the divisor unscaled, next to ordinary ones in
    the same waves, without any intermediate of the step overflowing."""
    shape = p.full_shape()
    out = {}
    for f in (("u", "v", "w") if p.dim == 3 else ("u", "v")):
        a = rng.choice([-1.0, 1.0], size=shape) * 10.0 ** rng.uniform(-305.0, 120.0, size=shape)
        r = rng.random(shape)
        a[r < 0.1] = 0.0
        a[(r >= 0.1) & (r < 0.15)] = -0.0
        band = a[..., : shape[-1] // 3]
        band[...] = rng.uniform(-2.0, 2.0, size=band.shape)
        out[f] = a
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_gpu_predictor_extreme_dividends_match_oracle(dim):
    """The predictor divides by the grid spacings and squared wall thicknesses with the
    divisor's reciprocal prepared once per wave, and falls back to true division for a
    cell whose dividends make v_div_scale rescale the divisor.  One step from velocities
    that reach both ends of the double range must equal the oracle bit for bit."""
    import pyoracle
    if dim == 3:
        p, phi = cases.sphere_case(m=40, n=24, l=20, steps=1, iter_max=4)
    else:
        p, phi = cases.cylinder_case(m=96, n=48, steps=1, iter_max=4)
    print(f"[gpu] predictor extreme dividends dim={dim}", flush=True)
    vel = _extreme_velocities(p, np.random.default_rng(11 + dim))
    o = pyoracle.OracleRun(p, phi)
    for f, a in vel.items():
        {"u": o.u, "v": o.v, "w": o.w}[f][...] = a
    e_ref = o.step(1)
    with _solver(p, phi) as s:
        for f, a in vel.items():
            s.set_field(f, a)
        st = s.step(1)
        np.testing.assert_array_equal(st["sor_error"], e_ref)
        for f in _fields(p):
            a, b = s.field(f), {"u": o.u, "v": o.v, "w": o.w, "p": o.p_}[f]
            bad = np.count_nonzero(a.view(np.uint64) != b.view(np.uint64))
            assert bad == 0, f"{f}: {bad} cells differ in their bits"

def test_gpu_full_size_sphere_two_steps():
    """256x128x128 sphere, 2 steps of 100 SOR iterations: rel-L2 <= 1e-12 per field."""
    import pyoracle
    p, phi = cases.sphere_case(steps=2)
    print("[gpu] full-size 256x128x128 sphere, 2 steps", flush=True)
    o = pyoracle.OracleRun(p, phi)
    e_ref = o.step(2)
    with _solver(p, phi) as s:
        st = s.step(2)
        for f in ("u", "v", "w", "p"):
            a, b = s.interior(f), o.interior(f)
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
            assert rel <= 1e-12, (f, rel)
            assert np.array_equal(a, b), f"{f}: not bit-identical (rel-L2 {rel:.3e})"
        np.testing.assert_array_equal(st["sor_error"], e_ref)


def test_gpu_step_chunks_are_consistent():
    """n steps in one call == the same steps split over several calls."""
    p, phi = cases.sphere_case(m=32, n=16, l=16, steps=6, iter_max=20)
    with _solver(p, phi) as a, _solver(p, phi) as b:
        ea = a.step(6)["sor_error"]
        eb = np.concatenate([b.step(1)["sor_error"], b.step(2)["sor_error"], b.step(3)["sor_error"]])
        np.testing.assert_array_equal(ea, eb)
        for f in ("u", "v", "w", "p"):
            np.testing.assert_array_equal(a.field(f), b.field(f))


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("name", ["g2d_weird_odd", "g3d_sph32", "g3d_odd_33_17_15", "g3d_air_room24", "g3d_weird"])
def test_gpu_sweep_kernels_agree(name, variant, monkeypatch):
    """The three colour-sweep kernels (stored faces; faces from porosity with IEEE division;
    with the checked constant-divisor division) reproduce the reference identically."""
    monkeypatch.setenv("PF_SOR_VARIANT", str(variant))
    p, g = load_golden(name)
    with _solver(p, g["phi_raw"]) as s:
        assert s.sor_kernel() == variant
        st = s.step(p.istep_max)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), g[f], err_msg=f"{name}:{f}:variant{variant}")
        np.testing.assert_array_equal(st["sor_error"], g["errs"])


@pytest.mark.parametrize("coded", ["0", "1"])
@pytest.mark.parametrize("name", ["g2d_weird_odd", "g3d_sph32", "g3d_odd_33_17_15", "g3d_air_room24", "g3d_weird",
                                  "g3d_boxes"])
def test_gpu_sweep_porosity_storage_matches_reference(name, coded, monkeypatch):
    """The FP64 sweep with the porosity read as stored FP64 (PF_PHI_CODE=0, the default at
    these sizes) and coded in 4 bytes (PF_PHI_CODE=1) reproduces the reference identically; the binary-porosity boxes
    hold the clamped 1e-6, which FP32 cannot hold exactly, and the tanh interfaces hold
    values that need the table."""
    monkeypatch.setenv("PF_PHI_CODE", coded)
    p, g = load_golden(name)
    with _solver(p, g["phi_raw"]) as s:
        assert s.sor_coded() == (coded == "1")
        st = s.step(p.istep_max)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), g[f], err_msg=f"{name}:{f}:coded{coded}")
        np.testing.assert_array_equal(st["sor_error"], g["errs"])


def test_gpu_porosity_code_falls_back_past_its_table(monkeypatch):
    """A porosity field with more distinct non-FP32 values than the code's table holds
    (2^20) is read as stored FP64: 120x96x100 cells of random porosity, checked against the
    oracle for one step."""
    import pyoracle
    monkeypatch.setenv("PF_PHI_CODE", "1")
    p = cases.CaseParams(dim=3, m=120, n=96, l=100, width=1.0, height=0.8, depth=0.83, time=1e-4, istep_max=1,
                         iter_max=5)
    phi = np.random.default_rng(7).uniform(0.2, 1.0, size=(p.l, p.n, p.m))
    with _solver(p, phi) as s:
        assert not s.sor_coded()
        s.step(1)
        o = pyoracle.OracleRun(p, phi)
        o.step(1)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), o.interior(f), err_msg=f)


def test_gpu_default_sweep_is_checked_fast_path():
    p, phi = cases.sphere_case(m=64, n=32, l=32, steps=1)
    with _solver(p, phi) as s:
        assert s.sor_kernel() == 2


def _force_scale(p, phi, pp, u, v, w):
    """Sum over cells of |term| for each of the 6 force components (numpy, any order): the
    scale that bounds the rounding difference between two summation orders."""
    sl = (slice(1, -1),) * p.dim
    def d(a, ax):
        return (np.roll(a, -1, axis=ax) - np.roll(a, 1, axis=ax))[sl] * 0.5
    axes = (1, 0) if p.dim == 2 else (2, 1, 0)
    g = [d(phi, ax) for ax in axes]
    na = np.sqrt(sum(x * x for x in g))
    den = np.maximum(na, 1e-6)
    ph = phi[sl]
    q = 1.0 - ph
    dv = p.dx * p.dy * (p.dz if p.dim == 3 else 1.0)
    hs = (p.dx, p.dy, p.dz)
    vel = (u, v, w)
    out = np.zeros(6)
    for c in range(p.dim):
        out[c] = np.abs(dv * pp[sl] * 2 * ph * q / (p.thickness * hs[c]) * g[c] / den).sum()
        t = ph * q / (p.thickness * hs[c])
        out[3 + c] = np.abs(dv * 32.0 * p.density * p.xnue * t * t * vel[c][sl]).sum()
    return out


def test_gpu_force_matches_reference_drag_log():
    """The drag executable's per-step force log (reference run on one thread): every step,
    to the rounding bound of a different summation order."""
    p, g = load_golden("g2d_drag")
    with _solver(p, g["phi_raw"]) as s:
        for st in range(p.istep_max):
            s.step(1)
            f = s.force()
            scale = _force_scale(p, s.field("porosity"), s.field("p"), s.field("u"), s.field("v"), None)
            ref = np.concatenate([g["force_p"][st], [0.0], g["force_v"][st], [0.0]])
            tol = 1e-13 * scale + 1e-300
            assert np.all(np.abs(f - ref) <= tol), (st, f, ref, tol)


def test_gpu_force_3d_matches_oracle():
    import pyoracle
    p, phi = cases.sphere_case(m=48, n=24, l=24, steps=3)
    o = pyoracle.OracleRun(p, phi)
    o.step(3)
    ref = o.force()
    with _solver(p, phi) as s:
        s.step(3)
        f = s.force()
        scale = _force_scale(p, s.field("porosity"), s.field("p"), s.field("u"), s.field("v"), s.field("w"))
    assert np.all(np.abs(f - ref) <= 1e-13 * scale), (f, ref, scale)
    assert abs(ref[0]) > 0 and abs(ref[3]) > 0


def test_gpu_boxes_512x256x256_one_step_matches_oracle():
    """The 512x256x256 boxes grid (40 seeded boxes, tanh voxel interface), one full step
    of 100 SOR iterations: bit-identical to the oracle."""
    import pyoracle
    p, phi = cases.boxes_case(steps=1)
    print("[gpu] full-size 512x256x256 boxes, 1 step", flush=True)
    o = pyoracle.OracleRun(p, phi)
    e_ref = o.step(1)
    with _solver(p, phi) as s:
        st = s.step(1)
        for f in ("u", "v", "w", "p"):
            np.testing.assert_array_equal(s.interior(f), o.interior(f), err_msg=f)
        np.testing.assert_array_equal(st["sor_error"], e_ref)
    o.close()


FUSED_CASES = ["g2d_cyl48", "g2d_c1_10", "g2d_backstep", "g3d_sph32", "g3d_stage_i1_s1", "g3d_weird", "g3d_air_room24",
               "g3d_boxes", "g3d_odd_32_17_15"]


@pytest.mark.parametrize("name", FUSED_CASES)
def test_gpu_fused_sweep_matches_reference(name, monkeypatch):
    """One launch per SOR iteration (both colours, ping-pong buffers) against the
    reference fixture, and against the two-launch sweep: the same bits."""
    p, g = load_golden(name)
    res = {}
    for fused in (1, 0):
        monkeypatch.setenv("PF_FUSED", str(fused))
        with _solver(p, g["phi_raw"]) as s:
            expect = fused == 1 and (p.bc == "air" or (p.n % 2 == 0 and (p.dim == 2 or p.l % 2 == 0)))
            assert s.sor_fused() == expect, name
            st = s.step(p.istep_max)
            res[fused] = ({f: s.interior(f) for f in _fields(p)}, st["sor_error"])
    for f in _fields(p):
        np.testing.assert_array_equal(res[1][0][f], g[f], err_msg=f"{name}:{f} fused")
        np.testing.assert_array_equal(res[1][0][f], res[0][0][f], err_msg=f"{name}:{f}")
    np.testing.assert_array_equal(res[1][1], g["errs"])


@pytest.mark.parametrize("name", ["3d_sphere_64x32x32", "3d_air_room_40x36x32", "2d_c1_256x128"])
def test_gpu_fused_mixed_matches_two_launch(name, monkeypatch):
    """FP32 fused iterations give the bits of the two-launch FP32 sweeps."""
    from pixelflow_amd.solver import PF_PREC_MIXED, Solver
    p, phi = CASES[name]()
    out = {}
    for fused in (1, 0):
        monkeypatch.setenv("PF_FUSED", str(fused))
        with Solver(p, phi, precision=PF_PREC_MIXED, inner_iters=10) as s:
            assert s.sor_fused() == bool(fused)
            st = s.step(p.istep_max)
            out[fused] = ({f: s.interior(f) for f in _fields(p)}, st["sor_error"])
    for f in _fields(p):
        np.testing.assert_array_equal(out[1][0][f], out[0][0][f], err_msg=f"{name}:{f}")
    np.testing.assert_array_equal(out[1][1], out[0][1])


@pytest.mark.parametrize("kb", [1, 2, 4])
@pytest.mark.parametrize("name", ["g3d_sph32", "g3d_odd_33_17_15", "g3d_air_room24", "g3d_weird", "g3d_boxes"])
def test_gpu_sweep_cells_per_thread_matches_reference(name, kb, monkeypatch):
    """3D colour sweeps with 1, 2 or 4 cells per thread (every cell's operands loaded before
    any cell is updated; a plane past the last one is loaded from plane l and skipped)
    reproduce the reference; odd l puts such planes in the last block row."""
    monkeypatch.setenv("PF_SWEEP_KB", str(kb))
    p, g = load_golden(name)
    with _solver(p, g["phi_raw"]) as s:
        st = s.step(p.istep_max)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), g[f], err_msg=f"{name}:{f}:kb{kb}")
        np.testing.assert_array_equal(st["sor_error"], g["errs"])


@pytest.mark.parametrize("shape", ["1,1", "1,5", "2,1", "2,3", "3,2", "4,8"])
@pytest.mark.parametrize("name", ["g2d_cyl48", "g2d_c1_10", "g2d_drag", "g2d_backstep", "g2d_stage_i1_s3",
                                  "g2d_weird_odd"])
def test_gpu_blocked_2d_sweep_matches_reference(name, shape, monkeypatch):
    """The temporally blocked 2D sweep (K iterations per launch in LDS, R rows per block,
    rows wrapped periodically in y; launches split so their count is even) reproduces the
    reference.  Shapes whose rows do not fit LDS (K = 3, 4 on the 256-wide grid) and the odd
    extent of g2d_weird_odd keep the two-launch sweeps."""
    monkeypatch.setenv("PF_SOR_TB2", shape)
    p, g = load_golden(name)
    print(f"[gpu] blocked 2D sweep {shape} fixture {name}", flush=True)
    with _solver(p, g["phi_raw"]) as s:
        st = s.step(p.istep_max)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), g[f], err_msg=f"{name}:{f}:tb2 {shape}")
        np.testing.assert_array_equal(st["sor_error"], g["errs"])
        np.testing.assert_array_equal(s.divergence(), g["div"], err_msg=f"{name}:div")


@pytest.mark.parametrize("name", ["3d_sphere_64x32x32", "3d_odd_45x27x19", "3d_air_room_40x36x32"])
def test_gpu_mixed_cells_per_thread_identical(name, monkeypatch):
    """FP32 sweeps give the same bits with 1, 2 or 4 cells per thread."""
    from pixelflow_amd.solver import PF_PREC_MIXED, Solver
    p, phi = CASES[name]()
    out = {}
    for kb in (1, 2, 4):
        monkeypatch.setenv("PF_SWEEP_KB", str(kb))
        with Solver(p, phi, precision=PF_PREC_MIXED, inner_iters=10) as s:
            st = s.step(p.istep_max)
            out[kb] = ({f: s.interior(f) for f in _fields(p)}, st["sor_error"])
    for kb in (2, 4):
        for f in _fields(p):
            np.testing.assert_array_equal(out[kb][0][f], out[1][0][f], err_msg=f"{name}:{f}:kb{kb}")
        np.testing.assert_array_equal(out[kb][1], out[1][1])


@pytest.mark.parametrize("name", ["g3d_sph32", "g3d_odd_33_17_15", "g3d_air_room24", "g3d_boxes", "g2d_cyl48",
                                  "g2d_weird_odd"])
@pytest.mark.parametrize("xw", [32, 64])
def test_gpu_sweep_wave_width_matches_reference(name, xw, monkeypatch):
    """Colour sweeps whose waves cover 64 slots of one row or 32 slots of two rows (the
    library picks the one that leaves fewer lanes idle) reproduce the reference."""
    monkeypatch.setenv("PF_SWEEP_XW", str(xw))
    p, g = load_golden(name)
    print(f"[gpu] wave width {xw} fixture {name}", flush=True)
    with _solver(p, g["phi_raw"]) as s:
        st = s.step(p.istep_max)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), g[f], err_msg=f"{name}:{f}:xw{xw}")
        np.testing.assert_array_equal(st["sor_error"], g["errs"])


@pytest.mark.parametrize("name", ["3d_sphere_64x32x32", "3d_odd_45x27x19", "2d_odd_97x51"])
def test_gpu_mixed_wave_width_identical(name, monkeypatch):
    """FP32 sweeps give the same bits with either wave width."""
    from pixelflow_amd.solver import PF_PREC_MIXED, Solver
    p, phi = CASES[name]()
    out = {}
    for xw in (32, 64):
        monkeypatch.setenv("PF_SWEEP_XW", str(xw))
        with Solver(p, phi, precision=PF_PREC_MIXED, inner_iters=10) as s:
            st = s.step(p.istep_max)
            out[xw] = ({f: s.interior(f) for f in _fields(p)}, st["sor_error"])
    for f in _fields(p):
        np.testing.assert_array_equal(out[32][0][f], out[64][0][f], err_msg=f"{name}:{f}")
    np.testing.assert_array_equal(out[32][1], out[64][1])


@pytest.mark.parametrize("remap", ["0", "1"])
@pytest.mark.parametrize("name", ["g3d_sph32", "g3d_odd_33_17_15", "g2d_cyl48"])
def test_gpu_block_order_matches_reference(name, remap, monkeypatch):
    """Sweeps, residuals and stencils in the hardware's block order and in the
    XCD-contiguous order (the library picks per precision and size) reproduce the
    reference: the order only changes which XCD runs a tile."""
    monkeypatch.setenv("PF_XCD_REMAP", remap)
    p, g = load_golden(name)
    print(f"[gpu] block order remap={remap} fixture {name}", flush=True)
    with _solver(p, g["phi_raw"]) as s:
        st = s.step(p.istep_max)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), g[f], err_msg=f"{name}:{f}:remap{remap}")
        np.testing.assert_array_equal(st["sor_error"], g["errs"])


@pytest.mark.parametrize("name", ["3d_sphere_64x32x32", "3d_air_room_40x36x32"])
def test_gpu_mixed_block_order_identical(name, monkeypatch):
    """FP32 sweeps and FP64 residuals give the same bits in either block order."""
    from pixelflow_amd.solver import PF_PREC_MIXED, Solver
    p, phi = CASES[name]()
    out = {}
    for remap in ("0", "1"):
        monkeypatch.setenv("PF_XCD_REMAP", remap)
        with Solver(p, phi, precision=PF_PREC_MIXED, inner_iters=10) as s:
            st = s.step(p.istep_max)
            out[remap] = ({f: s.interior(f) for f in _fields(p)}, st["sor_error"])
    for f in _fields(p):
        np.testing.assert_array_equal(out["1"][0][f], out["0"][0][f], err_msg=f"{name}:{f}")
    np.testing.assert_array_equal(out["1"][1], out["0"][1])


def _sparse_porosity(p, seed):
    """Porosity 1 with isolated cells of other values: on the domain's first and last rows,
    planes and columns (whose ghosts the periodic and outflow copies spread), on both
    sides of the 32-slot chunk boundaries (i = 63..66 and 127..130), and at random."""
    rng = np.random.default_rng(seed)
    shape = (p.l, p.n, p.m) if p.dim == 3 else (p.n, p.m)
    phi = np.ones(shape)
    if p.dim == 3:
        spots = [(0, 0, 0), (p.l - 1, p.n - 1, p.m - 1), (5, 0, 63), (7, p.n - 1, 64), (0, 17, 65),
                 (p.l - 1, 3, 66), (11, 20, 127), (3, 9, 128), (9, 30, 129), (2, 12, 1)]
    else:
        spots = [(0, 0), (p.n - 1, p.m - 1), (0, 63), (p.n - 1, 64), (17, 65), (3, 66), (20, 127), (9, 128),
                 (30, 129), (12, 1)]
    spots += [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(6)]
    for c in spots:
        phi[c] = rng.uniform(0.3, 0.95)
    return phi


@pytest.mark.parametrize("variant", ["default", "coded", "kb2", "xw64", "flags_off", "2d"])
def test_gpu_unit_porosity_waves_match_oracle(variant, monkeypatch):
    """FP64 colour sweeps skip the porosity loads of waves whose cells and neighbours all
    have porosity exactly 1 (and use 1 for those operands).  A field of porosity 1 with
    isolated other values at the faces, at the chunk boundaries and at random: a wave taken
    as unit where one of its operands is not 1 would change the pressure.  Three steps
    against the oracle, bit for bit, with each launch shape and porosity storage."""
    import pyoracle
    env = {"default": {}, "coded": {"PF_PHI_CODE": "1"}, "kb2": {"PF_SWEEP_KB": "2"}, "xw64": {"PF_SWEEP_XW": "64"},
           "flags_off": {"PF_UNIT_PHI": "0"}, "2d": {"PF_SOR_TB2": "0,0"}}[variant]
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    if variant == "2d":
        p = cases.CaseParams(dim=2, m=134, n=34, width=1.34, height=0.34, time=3e-4, istep_max=3, iter_max=20)
    else:
        p = cases.CaseParams(dim=3, m=134, n=34, l=22, width=1.34, height=0.34, depth=0.22, time=3e-4, istep_max=3,
                             iter_max=20)
    phi = _sparse_porosity(p, 11)
    print(f"[gpu] unit-porosity waves: {variant}", flush=True)
    o = pyoracle.OracleRun(p, phi)
    e_ref = o.step(p.istep_max)
    with _solver(p, phi) as s:
        unit = s.sor_unit_cells()
        if variant == "flags_off":
            assert unit == 0
        else:
            assert 0 < unit < p.cells
        st = s.step(p.istep_max)
        np.testing.assert_array_equal(st["sor_error"], e_ref)
        for f in _fields(p):
            np.testing.assert_array_equal(s.interior(f), o.interior(f), err_msg=f"{variant}:{f}")


def test_gpu_unit_porosity_cell_count_extremes(monkeypatch):
    """All cells in unit waves when the porosity is 1 everywhere (ghosts included; FP64 and
    the mixed-precision FP32 sweeps), none when no value is 1, none with the flags turned
    off."""
    p = cases.CaseParams(dim=3, m=64, n=32, l=16, width=1.0, height=0.5, depth=0.25, time=1e-4, istep_max=1,
                         iter_max=5)
    ones = np.ones((p.l, p.n, p.m))
    with _solver(p, ones) as s:
        assert s.sor_unit_cells() == p.cells
    rnd = np.random.default_rng(3).uniform(0.2, 0.99, size=ones.shape)
    with _solver(p, rnd) as s:
        assert s.sor_unit_cells() == 0
    from pixelflow_amd.solver import PF_PREC_MIXED, Solver
    with Solver(p, ones, precision=PF_PREC_MIXED, inner_iters=5) as s:   # the FP32 sweeps' flags
        assert s.sor_unit_cells() == p.cells
    monkeypatch.setenv("PF_UNIT_PHI", "0")
    with _solver(p, ones) as s:
        assert s.sor_unit_cells() == 0


@pytest.mark.parametrize("mode", ["mixed_fixed", "mixed_mg", "fp64_mg", "fp64_tol", "mixed_fixed_2d"])
def test_gpu_unit_porosity_solvers_identical(mode, monkeypatch):
    """The mixed-precision FP32 sweeps (flags of the FP32 porosity, whose unit values are
    those that round to 1), the FP64 residual (flags of the FP64 porosity) and the
    multigrid's level-0 kernels give the same bits with and without unit-porosity waves,
    on the sparse field above plus values within an FP32 rounding of 1."""
    from pixelflow_amd import solver as S
    if mode.endswith("_2d"):
        p = cases.CaseParams(dim=2, m=134, n=34, width=1.34, height=0.34, time=3e-4, istep_max=3, iter_max=20)
        phi = _sparse_porosity(p, 5)
        phi[10, 40] = 1.0 - 2.0 ** -30   # 1.0 in FP32, not in FP64
        phi[22, 99] = 1.0 - 2.0 ** -20
    else:
        p = cases.CaseParams(dim=3, m=134, n=34, l=22, width=1.34, height=0.34, depth=0.22, time=3e-4, istep_max=3,
                             iter_max=20)
        phi = _sparse_porosity(p, 5)
        phi[4, 10, 40] = 1.0 - 2.0 ** -30   # 1.0 in FP32, not in FP64
        phi[9, 22, 99] = 1.0 - 2.0 ** -20
    kw = {"mixed_fixed": dict(precision=S.PF_PREC_MIXED, inner_iters=10),
          "mixed_fixed_2d": dict(precision=S.PF_PREC_MIXED, inner_iters=10),
          "mixed_mg": dict(precision=S.PF_PREC_MIXED, residual_tol=1e-9, solver=S.PF_SOLVER_MG, max_cycles=30),
          "fp64_mg": dict(precision=S.PF_PREC_FP64, residual_tol=1e-9, solver=S.PF_SOLVER_MG, max_cycles=30),
          "fp64_tol": dict(precision=S.PF_PREC_FP64, residual_tol=1e-7, inner_iters=10, max_cycles=4)}[mode]
    print(f"[gpu] unit porosity, {mode}", flush=True)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("PF_UNIT_PHI", flag)
        with S.Solver(p, phi, **kw) as s:
            st = s.step(p.istep_max)
            out[flag] = ({f: s.interior(f) for f in _fields(p)}, st)
    for f in _fields(p):
        np.testing.assert_array_equal(out["1"][0][f], out["0"][0][f], err_msg=f"{mode}:{f}")
    for key in ("sor_error", "residual", "sor_iters"):
        np.testing.assert_array_equal(out["1"][1][key], out["0"][1][key], err_msg=f"{mode}:{key}")
