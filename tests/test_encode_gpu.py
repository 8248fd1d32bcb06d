"""Byte-identical `.mars` output: GPU search table vs the reference CPU search.

The same host-side chain (walks, RD, format, rANS -- pinned to the
reference's stream hashes in test_codec_pins.py) encodes each plane twice:
once with the oracle's exhaustive search as provider, once with the HIP
search table (mars_hip_encode_plane).  The bytes must be equal.  The oracle
side precomputes its table with all host cores and serves the walk from it.
"""

import hashlib

import pytest

from mars_amd import synth
from mars_amd.codec import EncodeOptions, encode_plane, hip_encode_plane, read_mars, write_mars
from mars_amd.provider import EncodeParams, HipContext
from tests._providers import TableProvider, assert_tables_equal

pytestmark = pytest.mark.gpu


def _both(img, params, options, shifts, check_tables=True):
    """Encode with the oracle's table and with the GPU table.  With
    check_tables, every winner of the GPU table the encode read must also
    equal the oracle's (a wrong winner that does not change a split or a
    mode decision would leave the bytes alone)."""
    prov = TableProvider(img, params, shifts)
    ref, ref_st = encode_plane(img, params, options, prov.fn, None)
    ctx = HipContext(0)
    got, st = hip_encode_plane(ctx, img, params, options)
    if check_tables:
        n = assert_tables_equal(ctx.tables(), prov, f"{img.shape[1]}x{img.shape[0]}")
        print(f"  {n} table winners equal the oracle's", flush=True)
    ctx.close()
    # the stream the GPU path wrote reads back and re-writes to itself
    hdr, leaves = read_mars(got)
    assert write_mars(hdr, leaves) == got
    return ref, ref_st, got, st


def test_fixed_grid_256_threshold_walk():
    """256x256 seed-1 plane, fixed 8x8 ranges, stride 8, threshold walk."""
    img = synth.plane(256, 256, synth.WORKLOAD_SEEDS["fixed256"])
    p = EncodeParams(min_size=8, max_size=8, shift=8, t_rms=8.0, lambda_=None)
    ref, ref_st, got, st = _both(img, p, EncodeOptions(), [8])
    assert st.search_evals == ref_st.search_evals == 1024 * 961 * 8
    print(f"fixed256: {len(got)} bytes sha256 {hashlib.sha256(got).hexdigest()}", flush=True)
    assert got == ref


def test_quadtree_512_rd_lambda200():
    """512x512 seed-2 plane, quadtree 32->4, stride 4, RD at lambda 200, modes {0,2}."""
    img = synth.plane(512, 512, synth.WORKLOAD_SEEDS["quadtree512"])
    p = EncodeParams(min_size=4, max_size=32, shift=4, lambda_=200.0)
    o = EncodeOptions(allowed_modes=(True, False, True, False))
    ref, ref_st, got, st = _both(img, p, o, [4])
    assert (st.search_evals, st.warmup_evals) == (ref_st.search_evals, ref_st.warmup_evals)
    print(f"quadtree512: {len(got)} bytes sha256 {hashlib.sha256(got).hexdigest()} "
          f"leaves {st.n_leaves} modes {st.leaf_modes}", flush=True)
    assert got == ref


@pytest.mark.parametrize("lam,modes,density", [
    (50.0, (True, True, True, True), False),
    (800.0, (True, False, True, False), True),
])
def test_rd_variants_160x96(lam, modes, density):
    img = synth.plane(160, 96, 77)
    p = EncodeParams(min_size=4, max_size=16, shift=4, lambda_=lam)
    o = EncodeOptions(allowed_modes=modes, adaptive_density=density,
                      lambda_adaptive_qstep=density)
    ref, _, got, _ = _both(img, p, o, [4, 8] if density else [4])
    assert got == ref


@pytest.mark.parametrize("max_alfa,bits_alfa,bits_beta,zero,lam", [
    (0.5, 5, 6, 1, None),     # threshold walk, zero-alfa override on small codes
    (0.75, 4, 7, 0, 120.0),   # RD walk, a scale that is not a power of two
])
def test_quantiser_settings_odd_geometry(max_alfa, bits_alfa, bits_beta, zero, lam):
    """75x53 plane (border nodes down to size 2) with non-default quantisers."""
    img = synth.plane(75, 53, 91)
    p = EncodeParams(min_size=4, max_size=16, shift=4, max_alfa=max_alfa, bits_alfa=bits_alfa,
                     bits_beta=bits_beta, zero_threshold=zero, lambda_=lam)
    o = EncodeOptions(allowed_modes=(True, True, True, True)) if lam else EncodeOptions()
    print(f"encode 75x53 max_alfa {max_alfa} bits {bits_alfa}/{bits_beta} zero {zero} "
          f"lambda {lam}", flush=True)
    ref, ref_st, got, st = _both(img, p, o, [4])
    assert (st.search_evals, st.warmup_evals) == (ref_st.search_evals, ref_st.warmup_evals)
    assert got == ref


def test_cli_a_stream_from_the_gpu_table():
    """The reference's CLI-A stream (cli_a_gate.rs:55-56): the same encode with the
    GPU search table as provider gives the pinned MARC bytes."""
    import hashlib as H

    from mars_amd.codec import wrap_gray
    from tests.test_codec_pins import CLI_A_STREAM_SHA256, cli_a_gradient, cli_a_params

    img = cli_a_gradient()
    ctx = HipContext(0)
    data, _ = hip_encode_plane(ctx, img, cli_a_params(),
                               EncodeOptions(allowed_modes=(True, True, True, True)))
    ctx.close()
    assert H.sha256(wrap_gray(data)).hexdigest() == CLI_A_STREAM_SHA256
