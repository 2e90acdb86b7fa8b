"""sb_counts_noalloc (the entry point that allocates nothing) against sb_counts,
both called through the C ABI, on two calls in a row (the second smaller), in
FP32 and FP64, with mask auto / none / explicit and the F81 path; results must be
bitwise equal (the same kernels; per-column arithmetic does not depend on how many
waves run) and within the stated tolerances of the FP64 oracle on a column
sample.  The device's free memory must not change across the no-allocation calls,
and sizes beyond the scratch array's limits fail with SB_ERR_ARG."""
import ctypes
import sys

import numpy as np
import pytest
import torch

from oracle import subby_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(sb, leaves, C, seed, A=4):
    rng = np.random.default_rng(seed)
    if A == 4:
        m = sb.hky85_diag(2.0, rng.dirichlet(np.full(4, 5.0)))
    else:
        pi = rng.dirichlet(np.full(A, 3.0))
        m = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (A, A)), pi), pi)
    t = sb.synthetic.random_tree(leaves, rng, 0.05)
    aln = sb.synthetic.simulate(t, m, C, rng, gap_frac=0.03, ambig_frac=0.01)
    return t, m, aln


def _run(lib, fn, eng, mid, aln_d, parent, dist, dtype, f81, mode, mask_d, outs, stream):
    cnt, root, ll = outs
    return fn(eng.ctx, mid, ctypes.c_void_p(aln_d.data_ptr()), parent.ctypes.data, dist.ctypes.data,
              aln_d.shape[0], aln_d.shape[1], dtype, 1 if f81 else 0, mode,
              ctypes.c_void_p(mask_d.data_ptr()) if mask_d is not None else None,
              ctypes.c_void_p(cnt.data_ptr()), ctypes.c_void_p(root.data_ptr()),
              ctypes.c_void_p(ll.data_ptr()), ctypes.c_void_p(stream))


def test_noalloc_equals_counts():
    import subby_amd as sb
    from subby_amd import _lib
    eng = sb.Engine(0)
    lib = eng.lib
    dev = eng.device
    stream = torch.cuda.current_stream(dev).cuda_stream
    try:
        # two calls in a row, the second smaller
        for leaves, C, seed in ((64, 20000, 51), (20, 3001, 52)):
            t, m, aln = _case(sb, leaves, C, seed)
            R = aln.shape[0]
            aln_d = torch.from_numpy(aln).to(dev)
            parent = np.ascontiguousarray(t.parentIndex, dtype=np.int32)
            dist = np.ascontiguousarray(t.distanceToParent, dtype=np.float64)
            mid = eng.model_id(m)
            mk = torch.from_numpy((np.random.default_rng(seed).random((R, C)) < 0.7).astype(np.uint8)).to(dev)
            kept = {}
            for prec, tdt, dt in (('fp32', torch.float32, _lib.SB_F32), ('fp64', torch.float64, _lib.SB_F64)):
                for f81, mode in ((False, _lib.SB_MASK_AUTO), (False, _lib.SB_MASK_NONE),
                                  (False, _lib.SB_MASK_EXPLICIT), (True, _lib.SB_MASK_AUTO)):
                    say(f'noalloc case R={R} C={C} {prec} f81={f81} mask={mode}')
                    mask_d = mk if mode == _lib.SB_MASK_EXPLICIT else None
                    mk_outs = lambda: (torch.full((4, 4, C), float('nan'), dtype=tdt, device=dev),  # noqa: E731
                                       torch.full((4, C), float('nan'), dtype=tdt, device=dev),
                                       torch.full((C,), float('nan'), dtype=torch.float64, device=dev))
                    ref, got = mk_outs(), mk_outs()
                    _lib.check(_run(lib, lib.sb_counts, eng, mid, aln_d, parent, dist, dt, f81, mode,
                                    mask_d, ref, stream))
                    torch.cuda.synchronize()
                    free0 = torch.cuda.mem_get_info(dev)[0]
                    _lib.check(_run(lib, lib.sb_counts_noalloc, eng, mid, aln_d, parent, dist, dt, f81, mode,
                                    mask_d, got, stream))
                    torch.cuda.synchronize()
                    assert torch.cuda.mem_get_info(dev)[0] == free0
                    for a, b in zip(ref, got):
                        assert torch.equal(a, b)
                    if not f81 and mode == _lib.SB_MASK_AUTO:
                        kept[prec] = got
            # both calls also against the oracle on a column sample (eigensub path, mask auto)
            cols = np.arange(0, C, 7)[:400]
            rc, rl, rr = O.counts_loglike_root(aln[:, cols], O.Tree(*t), O.DiagModel(*m))
            for prec, tol in (('fp64', dict(c=1e-12, l=1e-12, r=1e-12)), ('fp32', dict(c=1e-4, l=1e-5, r=1e-5))):
                cnt, root, ll = (x.double().cpu().numpy() for x in kept[prec])
                e_c = np.max(np.abs(cnt[:, :, cols] - rc).max((0, 1)) / np.abs(rc).max((0, 1)))
                e_l = np.max(np.abs(ll[cols] - rl) / np.maximum(np.abs(rl), 1))
                e_r = np.abs(root[:, cols] - rr).max()
                say(f'   oracle {prec}: counts {e_c:.2e} logL {e_l:.2e} root {e_r:.2e}')
                assert e_c <= tol['c'] and e_l <= tol['l'] and e_r <= tol['r']
        # a large alphabet: the VALU large-alphabet kernel (the two-pass path's stage would
        # not fit), within the stated FP32 tolerances of the oracle and of sb_counts
        t, m, aln = _case(sb, 15, 203, 53, A=20)
        R, C = aln.shape
        aln_d = torch.from_numpy(aln).to(dev)
        parent = np.ascontiguousarray(t.parentIndex, dtype=np.int32)
        dist = np.ascontiguousarray(t.distanceToParent, dtype=np.float64)
        mid = eng.model_id(m)
        outs = (torch.empty((20, 20, C), device=dev), torch.empty((20, C), device=dev),
                torch.empty(C, dtype=torch.float64, device=dev))
        _lib.check(_run(lib, lib.sb_counts_noalloc, eng, mid, aln_d, parent, dist, _lib.SB_F32, False,
                        _lib.SB_MASK_AUTO, None, outs, stream))
        torch.cuda.synchronize()
        rc, rl, rr = O.counts_loglike_root(aln, O.Tree(*t), O.DiagModel(*m))
        got = outs[0].double().cpu().numpy()
        assert np.max(np.abs(got - rc).max((0, 1)) / np.abs(rc).max((0, 1))) <= 1e-4
        assert np.max(np.abs(outs[2].cpu().numpy() - rl) / np.maximum(np.abs(rl), 1)) <= 1e-5
        # beyond the limits: branch records of a 12,001-node FP64 A = 8 tree (14.2 MB) exceed the area
        rng = np.random.default_rng(54)
        big = sb.synthetic.random_tree(6001, rng, 0.05)
        pi8 = rng.dirichlet(np.full(8, 3.0))
        m8 = sb.diagonalize_rate_matrix(sb.reversible_rate_matrix(rng.lognormal(0, 1, (8, 8)), pi8), pi8)
        R = len(big.parentIndex)
        aln_d = torch.zeros((R, 64), dtype=torch.int8, device=dev)
        outs = (torch.empty((8, 8, 64), dtype=torch.float64, device=dev),
                torch.empty((8, 64), dtype=torch.float64, device=dev),
                torch.empty(64, dtype=torch.float64, device=dev))
        parent = np.ascontiguousarray(big.parentIndex, dtype=np.int32)
        dist = np.ascontiguousarray(big.distanceToParent, dtype=np.float64)
        rc = _run(lib, lib.sb_counts_noalloc, eng, eng.model_id(m8), aln_d, parent, dist, _lib.SB_F64, False,
                  _lib.SB_MASK_AUTO, None, outs, stream)
        assert rc == -1 and b'no-allocation' in lib.sb_last_error()
    finally:
        torch.cuda.synchronize()
        eng.close()
