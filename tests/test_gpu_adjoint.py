"""Adjoint gradients on the GPU against the oracle's reverse-mode transpose
(itself pinned by central finite differences in test_oracle.py) and against
central differences of the GPU forward."""
import numpy as np
import pytest

from helpers import gpu_context, oracle_problem, rel_l2, tiny_problem

pytestmark = pytest.mark.gpu


def _loss_grad(f, n, target):
    f.forward_record(n, 1)
    ir = f.receivers(True)
    return float(np.sum((ir - target) ** 2)), 2 * (ir - target)


@pytest.mark.parametrize("sponge", [0, 3])
def test_fp64_adjoint_matches_oracle(gpu, sponge):
    p = tiny_problem(21, dims=(20, 16, 12), steps=90, sponge=sponge)
    n = p["steps"]
    target = np.random.default_rng(1).standard_normal((len(p["rcv_idx"]), n)) * 1e-3
    f = gpu_context(p, "fp64")
    f.forward_record(n, 5)
    ir = f.receivers(True)
    ref_ir = oracle_problem(p).run(n, "f64")["ir"]
    np.testing.assert_array_equal(ir, ref_ir)
    dl = 2 * (ir - target)
    g = f.adjoint(dl.astype(np.float32))
    g_or = oracle_problem(p).adjoint(n, dl.astype(np.float32).astype(np.float64))
    print(f"fp64 adjoint vs oracle rel {rel_l2(g, g_or):.2e}", flush=True)
    assert rel_l2(g, g_or) <= 1e-10
    f.close()


@pytest.mark.parametrize("nmat", [3, 5])
def test_fp64_adjoint_material_counts(gpu, nmat):
    """Up to three materials the wave sums of the gradient go through one
    tree over 16 slots, with more through per-material sums: both against
    the oracle."""
    p = tiny_problem(23, dims=(20, 16, 12), steps=60, nmat=nmat)
    n = p["steps"]
    print(f"fp64 adjoint, {nmat} materials", flush=True)
    f = gpu_context(p, "fp64")
    f.forward_record(n, 3)
    ir = f.receivers(True)
    target = np.random.default_rng(7).standard_normal(ir.shape) * 1e-3
    dl = (2 * (ir - target)).astype(np.float32)
    g = f.adjoint(dl)
    g_or = oracle_problem(p).adjoint(n, dl.astype(np.float64))
    assert g.shape == g_or.shape == (nmat, 5)
    print(f"fp64 adjoint ({nmat} materials) vs oracle rel {rel_l2(g, g_or):.2e}", flush=True)
    assert rel_l2(g, g_or) <= 1e-10
    f.close()


def test_fp64_adjoint_receivers_sharing_cells(gpu):
    """Receivers that share a cell, off the list (three on one cell) and on
    it (two): each receiver's dL/dIR enters the adjoint field once, as the
    oracle's in-order adds do."""
    p = tiny_problem(22, dims=(20, 16, 12), steps=70, sponge=3)
    r = p["rcv_idx"]
    bidx = p["bidx"].astype(np.int64)
    on_list = int(bidx[len(bidx) // 2])
    p["rcv_idx"] = np.array([r[0], r[1], r[0], on_list, r[2], on_list, r[0]], np.int64)
    n = p["steps"]
    _say = lambda m: print(m, flush=True)
    _say("fp64 adjoint, receivers sharing cells")
    f = gpu_context(p, "fp64")
    f.forward_record(n, 4)
    ir = f.receivers(True)
    np.testing.assert_array_equal(ir[2], ir[0])
    np.testing.assert_array_equal(ir[5], ir[3])
    target = np.random.default_rng(5).standard_normal(ir.shape) * 1e-3
    dl = (2 * (ir - target)).astype(np.float32)
    g = f.adjoint(dl)
    g_or = oracle_problem(p).adjoint(n, dl.astype(np.float64))
    print(f"fp64 adjoint (shared receiver cells) vs oracle rel {rel_l2(g, g_or):.2e}", flush=True)
    assert rel_l2(g, g_or) <= 1e-10
    f.close()


def test_adjoint_after_the_boundary_list_grows(gpu):
    """Record and adjoint on every other list entry, then the whole list
    uploaded into the same context: the checkpointed filter states, traces
    and reverse filter states follow the new length, the stale recording is
    refused, and the new gradient equals a fresh context's bit for bit.  New
    materials drop the recording, and a ninth material is refused by the
    recording forward."""
    from virs_amd.fdtd import FDTDError
    p = tiny_problem(26, dims=(20, 16, 12), steps=60, sponge=2)
    n = p["steps"]
    half = dict(p, bidx=p["bidx"][::2], bmat=p["bmat"][::2])
    dl = np.random.default_rng(8).standard_normal((len(p["rcv_idx"]), n)).astype(np.float32)
    print(f"fp64 adjoint, list of {len(half['bidx'])} then {len(p['bidx'])} entries", flush=True)
    f = gpu_context(half, "fp64")
    f.forward_record(n, 4)
    g_half = f.adjoint(dl)
    f.upload_boundary(p["bidx"], p["bmat"])
    with pytest.raises(FDTDError, match="forward_record first"):
        f.adjoint(dl)
    f.start(n)
    f.forward_record(n, 4)
    g = f.adjoint(dl)
    fresh = gpu_context(p, "fp64")
    fresh.forward_record(n, 4)
    np.testing.assert_array_equal(g, fresh.adjoint(dl))
    g_or = oracle_problem(p).adjoint(n, dl.astype(np.float64))
    print(f"after the re-upload vs oracle rel {rel_l2(g, g_or):.2e}", flush=True)
    assert rel_l2(g, g_or) <= 1e-10
    assert not np.array_equal(g, g_half)
    f.forward_record(n, 4)
    f.set_materials(np.tile(p["coeffs"][:1], (9, 1)))
    with pytest.raises(FDTDError, match="forward_record first"):
        f.adjoint(dl)
    f.start(n)
    with pytest.raises(FDTDError, match="up to 8 materials"):
        f.forward_record(n, 4)
    f.close()
    fresh.close()


def test_fp32_adjoint_within_1e4_of_fp64(gpu):
    p = tiny_problem(22, dims=(22, 16, 14), steps=100, sponge=2)
    n = p["steps"]
    target = np.random.default_rng(2).standard_normal((len(p["rcv_idx"]), n)) * 1e-3
    f64 = gpu_context(p, "fp64")
    f64.forward_record(n, 4)
    dl = (2 * (f64.receivers(True) - target)).astype(np.float32)
    g64 = f64.adjoint(dl)
    f32 = gpu_context(p, "fp32")
    f32.forward_record(n, 4)
    g32 = f32.adjoint(dl)
    e = rel_l2(g32, g64)
    print(f"fp32 adjoint vs fp64 rel {e:.2e}", flush=True)
    assert e <= 1e-4
    f64.close()
    f32.close()


@pytest.mark.parametrize("ncp", [0, 1, 2, 7, 100])
def test_checkpoint_count_does_not_change_the_gradient(gpu, ncp):
    """The whole trace (one segment), any number of checkpoints, and the
    automatic choice (0) give the same gradient bit for bit, also when the
    trace buffers are reused from a run with another segment length."""
    p = tiny_problem(23, dims=(18, 14, 12), steps=100)
    n = p["steps"]
    dl = np.random.default_rng(3).standard_normal((len(p["rcv_idx"]), n)).astype(np.float32)
    f = gpu_context(p, "fp64")
    f.forward_record(n, 1)
    ref = f.adjoint(dl)
    f.forward_record(n, ncp)
    g = f.adjoint(dl)
    np.testing.assert_array_equal(g, ref)
    f.forward_record(n, 1)
    np.testing.assert_array_equal(f.adjoint(dl), ref)
    f.close()


def test_fp64_adjoint_matches_central_differences(gpu):
    p = tiny_problem(24, dims=(16, 14, 12), steps=70, sponge=2)
    n = p["steps"]
    target = np.random.default_rng(4).standard_normal((len(p["rcv_idx"]), n)) * 1e-3
    f = gpu_context(p, "fp64")
    L0, dl = _loss_grad(f, n, target)
    f.forward_record(n, 3)
    g = f.adjoint(dl.astype(np.float32))
    c0 = p["coeffs"].copy()
    h = 1e-3
    fd = np.zeros_like(c0)
    for m in range(c0.shape[0]):
        for k in range(5):
            vals = []
            for sgn in (1, -1):
                c = c0.copy()
                c[m, k] += sgn * h
                f.set_materials(c)
                vals.append(_loss_grad(f, n, target)[0])
            fd[m, k] = (vals[0] - vals[1]) / (2 * h)
    f.set_materials(c0)
    e = rel_l2(g, fd)
    print(f"fp64 adjoint vs central differences rel {e:.2e}", flush=True)
    # dl was rounded to float32 for the adjoint call
    assert e <= 1e-5
    f.close()


def test_full_size_adjoint_against_central_differences_and_fp64(gpu):
    """The 256x192x128 shoebox-generator room, 3 materials (15 coefficients),
    4 receivers, 4096 steps, 64 checkpoints; target IRs from hidden seeded
    coefficients, L2 loss.  FP64 adjoint vs central differences (h = 1e-3,
    30 forward runs), FP32 adjoint vs FP64 adjoint."""
    from virs_amd import rooms
    from virs_amd.fdtd import setup
    T, ncp = 4096, 64
    sc = rooms.shoebox_grid((256, 192, 128), nmat=3, steps=T, nrcv=4, seed=7)
    prep = rooms.prepare(sc)
    rng = np.random.default_rng(99)
    target_c = np.array([rooms.random_biquad(rng) for _ in range(3)])
    grads = {}
    for prec in ("fp64", "fp32"):
        print(f"full-size adjoint {prec}", flush=True)
        f = setup(prep, prec)
        f.set_materials(target_c)
        f.start(T)
        f.step(T)
        target = f.receivers(True)
        f.set_materials(sc.coeffs)
        f.start(T)
        f.forward_record(T, ncp)
        grads[prec] = f.adjoint((2 * (f.receivers(True) - target)).astype(np.float32))
        if prec == "fp64":
            c0 = sc.coeffs.copy()
            fd = np.zeros_like(c0)
            for m in range(3):
                for k in range(5):
                    vals = []
                    for sgn in (1, -1):
                        c = c0.copy()
                        c[m, k] += sgn * 1e-3
                        f.set_materials(c)
                        f.start(T)
                        f.step(T)
                        vals.append(float(np.sum((f.receivers(True) - target) ** 2)))
                    fd[m, k] = (vals[0] - vals[1]) / 2e-3
            e_fd = rel_l2(grads["fp64"], fd)
            print(f"fp64 adjoint vs central differences rel {e_fd:.2e}", flush=True)
            assert e_fd <= 1e-5
        f.close()
    e = rel_l2(grads["fp32"], grads["fp64"])
    print(f"fp32 adjoint vs fp64 rel {e:.2e}", flush=True)
    assert e <= 1e-4


def test_torch_autograd_wrapper_matches_fd_and_fp64(gpu):
    """The adjoint behind torch.autograd: loss.backward() through BiquadIR
    gives the FP64 adjoint gradient and agrees with central differences."""
    import torch
    from torch_adjoint import BiquadIR
    p = tiny_problem(25, dims=(18, 16, 12), steps=80, sponge=2)
    n = p["steps"]
    target = torch.from_numpy(np.random.default_rng(5).standard_normal((len(p["rcv_idx"]), n)) * 1e-3)
    f = gpu_context(p, "fp64")
    c = torch.tensor(p["coeffs"], dtype=torch.float64, requires_grad=True)
    ir = BiquadIR.apply(c, f, n, 4)
    loss = ((ir - target) ** 2).sum()
    loss.backward()
    g = c.grad.numpy()
    # the same gradient straight from the C ABI
    f.set_materials(p["coeffs"])
    f.start(n)
    f.forward_record(n, 4)
    dl = 2 * (f.receivers(True) - target.numpy())
    g_direct = f.adjoint(dl.astype(np.float32))
    np.testing.assert_array_equal(g, g_direct)
    # central differences through the same wrapper's forward
    fd = np.zeros_like(p["coeffs"])
    with torch.no_grad():
        for m in range(fd.shape[0]):
            for k in range(5):
                vals = []
                for sgn in (1, -1):
                    cc = torch.tensor(p["coeffs"], dtype=torch.float64)
                    cc[m, k] += sgn * 1e-3
                    vals.append(float(((BiquadIR.apply(cc, f, n, 4) - target) ** 2).sum()))
                fd[m, k] = (vals[0] - vals[1]) / 2e-3
    e = rel_l2(g, fd)
    print(f"torch wrapper: adjoint vs central differences rel {e:.2e}", flush=True)
    assert e <= 1e-5
    f.close()
