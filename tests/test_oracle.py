"""Pins the CPU oracle (oracle/fdtd_oracle.c) against what the mathematics
fixes, since the reference holds no golden vectors or tests for this path.

Each check is chosen so a plausible slip in the restatement fails it:
  eigenmodes      wrong wall-mirror rule, wrong Courant constant, wrong sign
  energy          asymmetric neighbour handling
  reciprocity     a boundary correction that couples cells
  lfilter         filter recursion order / coefficient order / sign convention,
                  in both forms (float direct form I, double direct form II)
  single cell     the 0.1/0.9/0.4/K correction constants
  numpy float32   operand order and indexing of the FP32 dense update
  finite diffs    the hand-written adjoint
  RMS rule        the early-stop thresholds and smoothing
"""
import numpy as np
import pytest
from scipy.signal import lfilter

from helpers import oracle_problem, rel_l2, tiny_problem
from oracle import oracle as O


def walled_box(nx, ny, nz):
    f = np.zeros((nz, ny, nx), np.uint8)
    f[0, :, :] = f[-1, :, :] = 1
    f[:, 0, :] = f[:, -1, :] = 1
    f[:, :, 0] = f[:, :, -1] = 1
    return f


@pytest.mark.parametrize("modes", [(1, 0, 0), (2, 1, 0), (1, 2, 3), (5, 4, 2)])
def test_rigid_box_eigenmode(modes):
    nx, ny, nz = 14, 11, 9
    Lx, Ly, Lz = nx - 2, ny - 2, nz - 2
    kx, ky, kz = modes
    lam2 = 1.0 / 3.0
    h = 0.05
    dt = O.dt_for_lambda2(lam2, h)
    pb0 = O.Problem(walled_box(nx, ny, nz), h, dt=dt)
    lam2_f = float(pb0.c2dt2) * float(pb0.ih2)
    ii = np.arange(nx) - 1 + 0.5
    jj = np.arange(ny) - 1 + 0.5
    kk = np.arange(nz) - 1 + 0.5
    phi = (np.cos(np.pi * kz * kk / Lz)[:, None, None] * np.cos(np.pi * ky * jj / Ly)[None, :, None]
           * np.cos(np.pi * kx * ii / Lx)[None, None, :])
    phi[0, :, :] = phi[-1, :, :] = 0
    phi[:, 0, :] = phi[:, -1, :] = 0
    phi[:, :, 0] = phi[:, :, -1] = 0
    mu = 4 * sum(np.sin(np.pi * k / (2 * L)) ** 2 for k, L in ((kx, Lx), (ky, Ly), (kz, Lz)))
    w = np.arccos(1 - lam2_f * mu / 2)
    rc = (4, 5, 6)
    ridx = (rc[2] * ny + rc[1]) * nx + rc[0]
    pb = O.Problem(walled_box(nx, ny, nz), h, dt=dt, init_p=phi.reshape(-1),
                   init_q=(np.cos(w) * phi).reshape(-1), rcv_idx=[ridx])
    n = 200
    out = pb.run(n, "f64", want_fields=True)
    np.testing.assert_allclose(out["p"], np.cos(w * n) * phi, atol=1e-10)
    np.testing.assert_allclose(out["ir"][0], np.cos(w * np.arange(1, n + 1)) * phi.reshape(-1)[ridx],
                               atol=1e-10)


def _edge_energy(u1, u0, flags, lam2):
    """||u1-u0||^2 + lam2 * sum over air-air face pairs of (du1)(du0)."""
    air = flags != 1
    e = np.sum((u1 - u0)[air] ** 2)
    for ax in range(3):
        a = np.swapaxes(air, 0, ax)
        a1 = np.swapaxes(u1, 0, ax)
        a0 = np.swapaxes(u0, 0, ax)
        pair = a[1:] & a[:-1]
        e += lam2 * np.sum(((a1[1:] - a1[:-1]) * (a0[1:] - a0[:-1]))[pair])
    return e


def test_energy_conserved_with_walls():
    rng = np.random.default_rng(3)
    nx, ny, nz = 16, 13, 11
    f = walled_box(nx, ny, nz)
    f[1:-1, 1:-1, 1:-1][rng.random((nz - 2, ny - 2, nx - 2)) < 0.08] = 1
    p0 = rng.standard_normal(f.shape) * (f != 1)
    q0 = rng.standard_normal(f.shape) * (f != 1)
    h, lam2 = 0.05, 0.3
    pb = O.Problem(f, h, dt=O.dt_for_lambda2(lam2, h), init_p=p0.reshape(-1), init_q=q0.reshape(-1))
    l2 = float(pb.c2dt2) * float(pb.ih2)
    e0 = _edge_energy(p0, q0, f, l2)
    for n in (1, 17, 150):
        out = pb.run(n, "f64", want_fields=True)
        e = _edge_energy(out["p"], out["q"], f, l2)
        assert abs(e - e0) <= 1e-10 * abs(e0)


def test_reciprocity_with_boundary_filters():
    p = tiny_problem(11, sponge=0, hard=False, steps=150)
    a, b = int(p["soft_idx"][0]), int(p["rcv_idx"][0])
    sig = p["soft_sig"][:1]
    ab = O.Problem(p["flags"], p["h"], dt=p["dt"], bidx=p["bidx"], bmat=p["bmat"], coeffs=p["coeffs"],
                   soft_idx=[a], soft_sig=sig, rcv_idx=[b]).run(150, "f64")["ir"][0]
    ba = O.Problem(p["flags"], p["h"], dt=p["dt"], bidx=p["bidx"], bmat=p["bmat"], coeffs=p["coeffs"],
                   soft_idx=[b], soft_sig=sig, rcv_idx=[a]).run(150, "f64")["ir"][0]
    assert np.max(np.abs(ab)) > 1e-6
    np.testing.assert_allclose(ab, ba, rtol=0, atol=1e-12 * np.max(np.abs(ab)))


def _resonant_row(r=0.9995, theta=0.3, peak=0.95):
    """A band-pass biquad with poles at radius r, scaled to a peak gain."""
    from scipy.signal import freqz
    a1, a2 = -2.0 * r * np.cos(theta), r * r
    _, h = freqz([1.0, 0.0, -1.0], [1.0, a1, a2], worN=8192)
    g = peak / np.max(np.abs(h))
    return np.array([g, 0.0, -g, a1, a2])


@pytest.mark.parametrize("state_f32", [False, True])
@pytest.mark.parametrize("row", ["random", "rbj", "resonant"])
def test_boundary_filter_matches_lfilter(state_f32, row):
    """Every list entry's filter output over the run equals scipy's lfilter
    of its input sequence: the reference's float direct form I (state_f32)
    and the double direct form II of the FP32/FP64 GPU path.  rbj: the
    reference's default high-pass row (440 Hz, Q 0.1, simulator.cpp:955-967),
    whose real pole at 0.99909 makes the form II state about 1e4 times its
    input at low frequencies; resonant: poles at radius 0.9995, where the
    state exceeds the input by about 3e3 near resonance."""
    rbj = row != "random"
    p = tiny_problem(12, steps=400 if rbj else 80)
    if row == "rbj":
        from virs_amd import grid as G
        r = G.rbj_highpass(88200.0 * 2.0 * np.sqrt(3.0), 440.0, 0.1)
        p["coeffs"] = np.tile(np.asarray(r, np.float64), (len(p["coeffs"]), 1))
    elif row == "resonant":
        p["coeffs"] = np.tile(_resonant_row(), (len(p["coeffs"]), 1))
    nst = p["steps"]
    out = oracle_problem(p, state_f32=state_f32).run(nst, "f64" if not state_f32 else "f32",
                                                   want_trace=True)
    td, ty = out["trace_d"], out["trace_y"]
    tol = 1e-12 if not state_f32 else 2e-5
    for e in range(0, len(p["bidx"]), 37):
        c = p["coeffs"][p["bmat"][e]]
        if state_f32:
            c = c.astype(np.float32).astype(np.float64)
        y = lfilter(c[:3], [1.0, c[3], c[4]], td[:, e])
        scale = max(np.max(np.abs(y)), 1e-30)
        assert np.max(np.abs(y - ty[:, e])) <= tol * scale


def test_single_cell_boundary_update():
    """One interior cell with one wall face neighbour and five pressure-release
    neighbours, hand-derived from simulator_gpu.cu:83-88 and :130-148:
      R = 2 pc - q + c2dt2 ((pc + 0*5 - 6 pc) ih2)
      R -= K (0.1 y + 0.9 d) 0.4f,  d = pc - q,  y = b0 d (zero state)."""
    f = np.zeros((3, 3, 3), np.uint8)
    f[1, 1, 2] = 1  # +x neighbour is a wall
    f[1, 1, 1] = 2
    ci = 13
    pc, q = 0.8125, 0.3
    c = np.array([[0.6, -0.2, 0.05, -0.3, 0.1]])
    p0 = np.zeros(27)
    q0 = np.zeros(27)
    p0[ci], q0[ci] = pc, q
    h = 0.04
    pb = O.Problem(f, h, dt=O.dt_for_lambda2(1 / 3, h), bidx=[ci], coeffs=c, rcv_idx=[ci],
                   init_p=p0, init_q=q0)
    got = pb.run(1, "f64")["ir"][0, 0]
    c2dt2, ih2, K = float(pb.c2dt2), float(pb.ih2), float(pb.kconst)
    dense = 2 * pc - q + c2dt2 * ((pc - 6 * pc) * ih2)
    d = pc - q
    y = 0.6 * d
    want = dense - K * (0.1 * y + 0.9 * d) * float(np.float32(0.4))
    assert got == pytest.approx(want, rel=1e-15, abs=0)


def _numpy_dense_f32(P, Q, flags, c2dt2, ih2):
    """Vectorised float32 restatement of the reference dense update."""
    R = np.zeros_like(P)
    f = flags
    pc = P[1:-1, 1:-1, 1:-1]

    def V(sl):
        return np.where(f[sl] == 1, pc, P[sl])

    s = (V((slice(1, -1), slice(1, -1), slice(2, None))) + V((slice(1, -1), slice(1, -1), slice(0, -2))))
    s = s + V((slice(1, -1), slice(2, None), slice(1, -1)))
    s = s + V((slice(1, -1), slice(0, -2), slice(1, -1)))
    s = s + V((slice(2, None), slice(1, -1), slice(1, -1)))
    s = s + V((slice(0, -2), slice(1, -1), slice(1, -1)))
    lap = (s - np.float32(6) * pc) * ih2
    upd = (np.float32(2) * pc - np.float32(1) * Q[1:-1, 1:-1, 1:-1]) + c2dt2 * lap
    fi = f[1:-1, 1:-1, 1:-1]
    R[1:-1, 1:-1, 1:-1] = np.where(fi == 1, np.float32(0), np.where((fi == 0) | (fi == 2), upd, np.float32(0)))
    return R


def test_float32_dense_matches_numpy():
    rng = np.random.default_rng(5)
    nx, ny, nz = 17, 12, 10
    f = np.zeros((nz, ny, nx), np.uint8)
    f[rng.random(f.shape) < 0.15] = 1
    f[rng.random(f.shape) < 0.1] = 2
    P = (rng.standard_normal(f.shape)).astype(np.float32)
    Q = (rng.standard_normal(f.shape)).astype(np.float32)
    for arr in (P, Q):
        arr[0], arr[-1], arr[:, 0], arr[:, -1], arr[:, :, 0], arr[:, :, -1] = 0, 0, 0, 0, 0, 0
    h = 0.04
    pb = O.Problem(f, h, dt=O.dt_for_lambda2(0.3, h), init_p=P.reshape(-1).astype(np.float64),
                   init_q=Q.reshape(-1).astype(np.float64))
    P1, Q1 = P, Q
    for n in (1, 2, 3):
        R = _numpy_dense_f32(P1, Q1, f, pb.c2dt2, pb.ih2)
        P1, Q1 = R, P1
        out = pb.run(n, "f32", want_fields=True)
        np.testing.assert_array_equal(out["p"], P1.astype(np.float64))


def test_adjoint_matches_central_differences():
    p = tiny_problem(21, dims=(14, 12, 10), steps=70, sponge=2)
    rng = np.random.default_rng(0)
    pb = oracle_problem(p)
    n = p["steps"]
    target = rng.standard_normal((len(p["rcv_idx"]), n)) * 1e-3

    def loss(coeffs):
        pb.coeffs = np.ascontiguousarray(coeffs)
        ir = pb.run(n, "f64")["ir"]
        return float(np.sum((ir - target) ** 2)), ir

    c0 = p["coeffs"].copy()
    L0, ir0 = loss(c0)
    pb.coeffs = np.ascontiguousarray(c0)
    g = pb.adjoint(n, 2 * (ir0 - target))
    hstep = 1e-3
    fd = np.zeros_like(c0)
    for m in range(c0.shape[0]):
        for k in range(5):
            cp = c0.copy()
            cm = c0.copy()
            cp[m, k] += hstep
            cm[m, k] -= hstep
            fd[m, k] = (loss(cp)[0] - loss(cm)[0]) / (2 * hstep)
    assert rel_l2(g, fd) < 1e-5


def test_rms_rule_closed_forms():
    # nothing can stop at or before step 3000
    assert O.rms_continue_sequence(np.zeros(3001)) == -1
    # silence in the 100-sample window: rms 0 < target at the first check
    ir = np.ones(3200, np.float32)
    ir[2900:] = 0
    assert O.rms_continue_sequence(ir) == 3001
    # a constant amplitude above the target never stops
    assert O.rms_continue_sequence(np.full(6000, 0.002, np.float32)) == -1
    # just below the target stops at the first check
    assert O.rms_continue_sequence(np.full(3100, 0.0008, np.float32)) == 3001


def test_sponge_single_cell_closed_form():
    """A lone interior cell walled on all six faces (L = 0) at edge distance 1
    of a 2-cell sponge: p^{n+1} = a p^n - b p^{n-1} with roots 1 and
    b = (1-s)/(1+s), s = sigma/8: p^n = A + B b^n."""
    f = np.ones((3, 3, 3), np.uint8)
    f[1, 1, 1] = 0
    p0 = np.zeros(27)
    p0[13] = 1.0
    sigma = 0.75
    pb = O.Problem(f, 0.04, dt=O.dt_for_lambda2(1 / 3, 0.04), init_p=p0, rcv_idx=[13],
                   sponge_width=2, sponge_sigma=sigma)
    a, b = float(pb.spg[0][1]), float(pb.spg[1][1])
    s = 0.5 * sigma * 0.25
    assert a == pytest.approx(2 / (1 + s), rel=1e-7)
    assert b == pytest.approx((1 - s) / (1 + s), rel=1e-7)
    ir = pb.run(40, "f64")["ir"][0]
    # p^0 = 1, p^-1 = 0: A + B = 1, A + B/b = 0
    B = b / (b - 1)
    A = 1 - B
    n = np.arange(1, 41)
    np.testing.assert_allclose(ir, A + B * b ** n, rtol=1e-12)


def test_reference_cpu_step_single_cell():
    """simulator_cpu.cpp:55-81 on one cell with pressure-release neighbours:
    p^{n+1} = (2 - g dt) p - (1 - g dt) q + c2dt2 (-6 p) ih2 in float32."""
    f = np.zeros((3, 3, 3), np.uint8)
    p0 = np.zeros((3, 3, 3), np.float32)
    p0[1, 1, 1] = 1.0
    h = np.float32(0.01)
    out = O.refcpu_steps(f, float(h), 3, p=p0)
    c = np.float32(343)
    dt = np.float32(0.5 * float(h) / float(c * np.sqrt(np.float32(3))))
    c2 = c * c * dt * dt
    ih2 = np.float32(1) / (h * h)
    damp = np.float32(5) * dt
    p, q = np.float32(1), np.float32(0)
    for _ in range(3):
        lap = (np.float32(0) - np.float32(6) * p) * ih2
        p, q = (np.float32(2) - damp) * p - (np.float32(1) - damp) * q + c2 * lap, p
    assert out[1, 1, 1] == p


# --- 16-bit storage runner -------------------------------------------------
#   fp16 nearest    wrong rounding constant / exponent rebias (numpy is RNE)
#   bf16 nearest    a truncation or a tie that does not go to even
#   stochastic      a biased or wrongly masked random add (the mean over
#                   every random value must reproduce the input exactly)
#   scale shift     a power-of-two scale applied to the wrong field or step

def _fp16_test_values(rng, n):
    mags = np.exp(rng.uniform(np.log(1e-9), np.log(7e4), n)).astype(np.float32)
    x = mags * rng.choice([-1, 1], n).astype(np.float32)
    # exact ties halfway between neighbouring halfs, normal and subnormal
    h = rng.integers(1, 0x7bff, 2000, dtype=np.uint16).view(np.float16).astype(np.float64)
    nxt = (h.astype(np.float16).view(np.uint16) + 1).view(np.float16).astype(np.float64)
    ties = ((h + nxt) / 2).astype(np.float32)
    return np.concatenate([x, ties, np.float32([0.0, -0.0, 65504.0, 65519.0, 65520.0, 1e6, np.inf, -np.inf])])


def test_fp16_nearest_matches_numpy():
    x = _fp16_test_values(np.random.default_rng(0), 200000)
    bits, back = O.round16(x, "fp16")
    with np.errstate(over="ignore"):
        want = x.astype(np.float16)
    np.testing.assert_array_equal(bits, want.view(np.uint16))
    np.testing.assert_array_equal(back, want.astype(np.float32))


def test_16bit_products_in_the_update_are_exact_in_fp32():
    """kw*pc (kw = 0..6) and 2*pc are exact in FP32 for every finite FP16
    value and every BF16 value below 2^100 (the per-field scale keeps stored
    values near 2^14).  The 16-bit stencil relies on this when it computes
    sum - kw*pc and 2*pc - q as fused multiply-adds: each then has the bits
    of the product and the difference rounded apart, as the oracle's runner
    computes them."""
    h = np.arange(65536, dtype=np.uint16).view(np.float16).astype(np.float64)
    h = h[np.isfinite(h)]
    b32 = (np.arange(65536, dtype=np.uint32) << 16).astype(np.uint32).view(np.float32)
    b = b32[np.isfinite(b32)].astype(np.float64)
    b = b[np.abs(b) < 2.0 ** 100]
    for v in (h, b):
        for kw in range(7):
            prod = kw * v
            np.testing.assert_array_equal(prod.astype(np.float32).astype(np.float64), prod)
        np.testing.assert_array_equal((2 * v).astype(np.float32).astype(np.float64), 2 * v)
    # a general FP32 value is not enough: the claim needs the short significand
    f = np.float64(np.float32(1.0) + np.float32(2.0 ** -23))
    assert np.float64(np.float32(3 * f)) != 3 * f


def test_fp16_every_pattern_round_trips():
    pat = np.arange(0x10000, dtype=np.uint32).astype(np.uint16)
    h = pat.view(np.float16)
    keep = np.isfinite(h)
    bits, back = O.round16(h[keep].astype(np.float32), "fp16")
    np.testing.assert_array_equal(bits, pat[keep])
    np.testing.assert_array_equal(back, h[keep].astype(np.float32))


def test_bf16_nearest_is_the_nearest_value_ties_to_even():
    rng = np.random.default_rng(1)
    u = rng.integers(0, 0x7f7f0000, 200000, dtype=np.uint32)
    u[:2000] = (u[:2000] & 0xffff0000) | 0x8000          # exact ties
    x = u.view(np.float32) * rng.choice([-1, 1], u.size).astype(np.float32)
    bits, back = O.round16(x, "bf16")
    lo = (x.view(np.uint32) & 0xffff0000).view(np.float32).astype(np.float64)
    hi = ((x.view(np.uint32) & 0xffff0000) + 0x10000).view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dlo, dhi = np.abs(xd - lo), np.abs(xd - hi)
    got = back.astype(np.float64)
    assert np.all(np.abs(xd - got) <= np.minimum(dlo, dhi))
    tie = dlo == dhi
    assert np.all((bits[tie] & 1) == 0)
    np.testing.assert_array_equal(back.view(np.uint32) & 0xffff, 0)


@pytest.mark.parametrize("fmt,nbits", [("fp16", 13), ("bf16", 16)])
def test_stochastic_rounding_is_exactly_unbiased(fmt, nbits):
    """Over every value of the random bits, stochastic rounding returns one of
    the two neighbouring representable values and its mean is the input
    exactly (normal range: a float32 has exactly `nbits` more mantissa bits)."""
    rng = np.random.default_rng(2)
    lo, hi = (2.0 ** -14, 6.0e4) if fmt == "fp16" else (1e-30, 1e30)
    x = (np.exp(rng.uniform(np.log(lo), np.log(hi), 64)) * rng.choice([-1, 1], 64)).astype(np.float32)
    r = np.arange(1 << nbits, dtype=np.uint32)
    for v in x:
        _, back = O.round16(np.full(r.size, v, np.float32), fmt, r)
        vals = np.unique(back)
        assert len(vals) <= 2
        if fmt == "fp16":
            near = v.astype(np.float16).astype(np.float32)
            assert near in vals
        assert np.sum(back.astype(np.float64)) / r.size == float(v)


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
@pytest.mark.parametrize("stochastic", [False, True])
@pytest.mark.parametrize("k", [-7, 5])
def test_16bit_runner_is_invariant_under_power_of_two_input_scaling(fmt, stochastic, k):
    """Scaling every input by 2^k must leave the stored patterns unchanged and
    shift each field exponent by -k: the scale is applied exactly once to
    each load and store."""
    p = tiny_problem(40, steps=50)
    n = np.arange(p["steps"])
    hard = np.where(n < 10, 10.0, 0.0)
    outs = []
    for sc in (1.0, 2.0 ** k):
        q = dict(p)
        q["soft_sig"] = p["soft_sig"] * sc
        pb = oracle_problem(q)
        pb.hard_sig = hard * sc
        outs.append(pb.run16(p["steps"], fmt, stochastic, want_fields=True))
    a, b = outs
    np.testing.assert_array_equal(b["exp"], a["exp"] - k)
    np.testing.assert_array_equal(b["ir"], a["ir"] * 2.0 ** k)
    np.testing.assert_array_equal(b["p"], a["p"] * 2.0 ** k)
    np.testing.assert_array_equal(b["q"], a["q"] * 2.0 ** k)


@pytest.mark.parametrize("fmt,bound", [("fp16", 2e-3), ("bf16", 3e-2)])
def test_16bit_runner_tracks_the_fp64_scheme(fmt, bound):
    p = tiny_problem(41, steps=120)
    ref = oracle_problem(p).run(p["steps"], "f64")["ir"]
    got = oracle_problem(p).run16(p["steps"], fmt, True)["ir"]
    assert rel_l2(got, ref) <= bound


@pytest.mark.parametrize("precision,state_f32", [("f32", False), ("f64", False), ("f32", True)])
def test_chunked_runs_equal_one_run(precision, state_f32):
    """A run continued in chunks, carrying the fields (init_p/init_q), the
    filter state (state_io) and the step offset (n0), gives the bits of one
    uninterrupted run."""
    p = tiny_problem(41, steps=60)
    whole = oracle_problem(p, state_f32=state_f32).run(60, precision, want_fields=True)
    pb = oracle_problem(p, state_f32=state_f32)
    pb.state_io = np.zeros(4 * max(1, len(p["bidx"])), np.float64)
    irs = []
    for n0 in (0, 25, 40):
        n = {0: 25, 25: 15, 40: 20}[n0]
        out = pb.run(n, precision, n0=n0, want_fields=True)
        irs.append(out["ir"])
        pb.init_p = np.ascontiguousarray(out["p"].reshape(-1), np.float64)
        pb.init_q = np.ascontiguousarray(out["q"].reshape(-1), np.float64)
    np.testing.assert_array_equal(np.concatenate(irs, axis=1), whole["ir"])
    np.testing.assert_array_equal(out["p"], whole["p"])
    np.testing.assert_array_equal(out["q"], whole["q"])


def test_16bit_runner_with_the_reference_high_pass_row():
    """The 16-bit scheme keeps its boundary filter in the transposed direct
    form with the two states stored in float.  With the reference's default
    high-pass row (440 Hz, Q 0.1, simulator.cpp:955-967), whose real pole at
    0.99909 gives the filter a low-frequency gain near 1e4, the FP16 runner
    stays as close to the FP64 scheme as with a benign random row (measured
    2.95e-3 against 2.16e-3 over 1500 steps): the stored states' rounding
    stays below the fields' storage noise."""
    from virs_amd import grid as G
    row = np.asarray(G.rbj_highpass(88200.0 * 2.0 * np.sqrt(3.0), 440.0, 0.1), np.float64)
    errs = {}
    for name in ("random", "rbj"):
        p = tiny_problem(47, dims=(40, 32, 28), steps=1500, hard=False, sponge=0)
        if name == "rbj":
            p["coeffs"] = np.tile(row, (len(p["coeffs"]), 1))
        o = oracle_problem(p)
        r64 = o.run(1500, "f64")
        r16 = o.run16(1500, "fp16", True)
        errs[name] = max(rel_l2(r16["ir"][r], r64["ir"][r]) for r in range(len(p["rcv_idx"])))
    assert errs["rbj"] <= 5e-3
    assert errs["rbj"] <= 2.0 * errs["random"]
