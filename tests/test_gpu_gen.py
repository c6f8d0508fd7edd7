"""GPU generators and preprocessing (kz_gpu_gen_*): the graphs built in HBM
are bit-for-bit the host generators' graphs (same CSR digest), symmetric and
degree-oriented, also when the radix sort runs in many chunks."""

import pytest

pytestmark = pytest.mark.gpu


def _cases(kz):
    yield "er_2^16_2^20", lambda o: kz.gpu_gen_erdos_renyi(1 << 16, 1 << 20, 1, o), \
        lambda: kz.gen_erdos_renyi(1 << 16, 1 << 20, 1)
    yield "rmat16", lambda o: kz.gpu_gen_rmat(16, 16, 0.57, 0.19, 0.19, 3, o), \
        lambda: kz.gen_rmat(16, 16, 0.57, 0.19, 0.19, 3)
    yield "rmat12_ef4", lambda o: kz.gpu_gen_rmat(12, 4, 0.45, 0.15, 0.15, 9, o), \
        lambda: kz.gen_rmat(12, 4, 0.45, 0.15, 0.15, 9)
    yield "chunglu_50k", lambda o: kz.gpu_gen_chung_lu(50000, 300000, 2.3, 3000, 5, o), \
        lambda: kz.gen_chung_lu(50000, 300000, 2.3, 3000, 5)
    yield "er_tiny", lambda o: kz.gpu_gen_erdos_renyi(5, 3, 2, o), lambda: kz.gen_erdos_renyi(5, 3, 2)


@pytest.mark.parametrize("chunk", [1 << 30, 1 << 18])
def test_gpu_generators_match_host(kzdev, chunk):
    kz = kzdev
    kz.set_sort_chunk(chunk)
    try:
        for name, gpu, host in _cases(kz):
            h = host()
            ho = h.oriented()
            for orient, want in ((False, h), (True, ho)):
                print(f"[case] {name} orient={orient} chunk={chunk}", flush=True)
                dg = gpu(orient)
                got = dg.download()
                assert (got.num_vertices, got.num_entries, got.max_degree) == \
                    (want.num_vertices, want.num_entries, want.max_degree), name
                assert got.oriented_flag == orient
                assert got.digest() == want.digest(), (name, orient)
                dg.release()
    finally:
        kz.set_sort_chunk(1 << 30)


def test_gpu_generated_graph_counts(kzdev):
    """A GPU-built graph is used directly by the counting paths."""
    kz = kzdev
    dg = kz.gpu_gen_rmat(14, 16, seed=2, orient=True)
    ref = kz.gen_rmat(14, 16, seed=2).oriented().to_device()
    assert dg.count_cliques(4) == ref.count_cliques(4)
    sym = kz.gpu_gen_chung_lu(20000, 100000, 2.3, 1000, 4)
    assert sym.motif_census() == kz.gen_chung_lu(20000, 100000, 2.3, 1000, 4).to_device().motif_census()
