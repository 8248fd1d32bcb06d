"""mars_hip_search_grid_noalloc: the fixed-grid search without allocation.

Called through the C ABI with device buffers (torch is only the device-memory
plumbing) and checked field for field against the tables of
mars_hip_prepare with min_size == max_size, on several calls in a row whose
later inputs are smaller (the shared scratch then holds the previous call's
contents), plus the refusal of a plane whose step exceeds the scratch.
"""

import ctypes as C

import numpy as np
import pytest

from mars_amd import _native as N
from mars_amd import synth
from mars_amd.provider import EncodeParams, HipContext
from tests._providers import candidate_key

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _noalloc(ctx, img, params):
    L = N.lib()
    h, w = img.shape
    s = params.max_size
    nr = (w // s) * (h // s)
    dev = torch.device("cuda", 0)
    d_px = torch.from_numpy(np.ascontiguousarray(img)).to(dev)
    rec = C.sizeof(N.CCandidate)
    d_out = torch.full((max(nr, 1) * rec,), 0xAB, dtype=torch.uint8, device=dev)
    d_found = torch.full((max(nr, 1),), 7, dtype=torch.uint8, device=dev)
    p = N.Params(params.min_size, params.max_size, params.shift, params.bits_alfa,
                 params.bits_beta, N.METHOD_EXHAUSTIVE, params.max_alfa)
    stream = torch.cuda.current_stream(dev)
    rc = L.mars_hip_search_grid_noalloc(ctx._ctx, C.c_void_p(d_px.data_ptr()), w, h,
                                        C.byref(p), C.c_void_p(d_out.data_ptr()),
                                        C.c_void_p(d_found.data_ptr()),
                                        C.c_void_p(stream.cuda_stream))
    if rc != N.OK:
        return rc, L.mars_hip_last_error(ctx._ctx).decode(), None
    torch.cuda.synchronize(dev)
    raw = d_out.cpu().numpy().tobytes()
    found = d_found.cpu().numpy()[:nr]
    cands = [N.CCandidate.from_buffer_copy(raw, i * rec) if found[i] else None
             for i in range(nr)]
    assert set(found.tolist()) <= {0, 1}
    return rc, "", cands


def _prepared(img, params):
    ctx = HipContext(0)
    ctx.prepare(img, params, [params.shift])
    by = {}
    for size, shift, rows, cols, cands, _ in ctx.tables():
        if size == params.max_size:
            by = {(r, c): cand for r, c, cand in zip(rows, cols, cands)}
    ctx.close()
    return by


CASES = [
    # (w, h, seed, size, shift, quantiser): later calls smaller than earlier ones
    (384, 256, 61, 16, 4, {}),                       # threshold pre-pass, coalesced order
    (160, 96, 62, 4, 2, {}),
    (100, 70, 63, 8, 3, dict(max_alfa=0.5, bits_alfa=5, bits_beta=6)),  # odd stride, border
    (64, 64, 64, 32, 2, {}),
    (20, 12, 65, 8, 4, {}),                          # no legal domain position
]


def test_noalloc_equals_prepare_on_consecutive_calls():
    ctx = HipContext(0)
    for w, h, seed, s, shift, q in CASES:
        img = synth.plane(w, h, seed)
        params = EncodeParams(min_size=s, max_size=s, shift=shift, **q)
        print(f"noalloc {w}x{h} size {s} stride {shift} {q}", flush=True)
        rc, msg, got = _noalloc(ctx, img, params)
        assert rc == N.OK, msg
        want = _prepared(img, params)
        gw = w // s
        assert len(got) == (w // s) * (h // s) == len(want)
        for i, cand in enumerate(got):
            r, c = (i // gw) * s, (i % gw) * s
            assert candidate_key(cand) == candidate_key(want[(r, c)]), \
                f"{w}x{h} size {s}: range ({r},{c})"
        if h < 2 * s:
            assert all(x is None for x in got)
    ctx.close()


@pytest.mark.parametrize("s,shift", [(2, 4), (4, 2), (8, 2), (16, 2), (32, 1)])
def test_noalloc_fits_the_documented_1024_planes(s, shift):
    """The header's statement of what fits the scratch: a 1024 x 1024 plane
    at sizes 4..32 with stride >= 2 and at size 2 with stride >= 4 (stride 1
    fits at size 32 only); every range finds a winner."""
    ctx = HipContext(0)
    img = synth.plane(1024, 1024, 67)
    print(f"noalloc 1024x1024 size {s} stride {shift}", flush=True)
    rc, msg, got = _noalloc(ctx, img, EncodeParams(min_size=s, max_size=s, shift=shift))
    assert rc == N.OK, msg
    assert len(got) == (1024 // s) ** 2 and all(x is not None for x in got)
    if s == 2:
        rc, msg, _ = _noalloc(ctx, img, EncodeParams(min_size=2, max_size=2, shift=2))
        assert rc == N.E_UNSUPPORTED and "scratch" in msg, msg
    ctx.close()


def test_noalloc_refuses_what_exceeds_the_scratch():
    ctx = HipContext(0)
    img = np.zeros((4096, 4096), np.uint8)
    rc, msg, _ = _noalloc(ctx, img, EncodeParams(min_size=4, max_size=4, shift=1))
    assert rc == N.E_UNSUPPORTED and "scratch" in msg, msg
    rc, msg, _ = _noalloc(ctx, img[:64, :64], EncodeParams(min_size=4, max_size=8, shift=4))
    assert rc == N.E_ARG, msg
    # the context still works after a refusal
    img = synth.plane(96, 64, 66)
    params = EncodeParams(min_size=8, max_size=8, shift=4)
    rc, msg, got = _noalloc(ctx, img, params)
    assert rc == N.OK, msg
    want = _prepared(img, params)
    assert [candidate_key(x) for x in got] == \
        [candidate_key(want[((i // 12) * 8, (i % 12) * 8)]) for i in range(len(got))]
    ctx.close()
