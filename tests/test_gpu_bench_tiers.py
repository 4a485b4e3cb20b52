"""The bench's roofline credits each kernel with the algorithmic bytes of the
rows it finishes, so the bench bins rows itself (bench.bin_caps and
bench.tiers).  These tests hold that binning to the library's: the caps an
untuned batch plan resolves (the same binning as a single call), and the
library's row counts per tier, for B narrower than 2^16 columns, up to 2^20
and wider, under the knobs that change the caps."""
import pytest

import bench

pytestmark = pytest.mark.gpu

KNOBS = ("med_cap", "lds_cfg", "wmed_cap", "wmed_engine", "wmed_bitmap")


@pytest.mark.parametrize("ncols", [5000, 100000, (1 << 20) + 7], ids=["N5k", "N100k", "N2^20+7"])
@pytest.mark.parametrize("knobs", [{}, {"wmed_bitmap": 0}, {"wmed_engine": 0}, {"lds_cfg": 2}, {"wmed_cap": 256},
                                   {"wmed_cap": 2048, "med_cap": 1024}],
                         ids=["default", "sort_ranks", "lds_wave", "lds_2048", "wcap256", "wcap2048_med1024"])
def test_bench_binning_matches_library(gpu, ncols, knobs):
    bm = gpu
    saved = {k: bm.get_param(k) for k in KNOBS}
    try:
        for k, v in knobs.items():
            bm.set_param(k, v)
        B = bm.gen.er_var(3000, ncols, 5, 40, 3)
        A = bm.gen.er_var(800, 3000, 1, 300, 4)  # rows of 5 to about 6,600 products
        dA, dB = bm.DeviceCSR.upload(A), bm.DeviceCSR.upload(B)
        try:
            med_cap, wcap = bench.bin_caps(bm, B.N)
            plan = bm.BatchPlan([(dA, dB)], autotune=False)
            try:
                info = plan.info()
            finally:
                plan.close()
            assert (info["med_cap"], info["wmed_cap"]) == (med_cap, wcap)
            L, ip = bench.row_stats(A, B)
            t = bench.tiers(L, ip, med_cap, wcap)
            counts = [int((t == 0).sum()), int((t == 1).sum()), int(((t == 2) | (t == 4)).sum()),
                      int((t == 3).sum())]
            assert info["rows_by_tier"] == counts
            assert min(counts[1:]) > 0 and int((t == 4).sum()) > 0
            bm.spgemm_dev(dA, dB).free()
            assert [int(x) for x in bm.last_stats()["rows_by_tier"]] == counts
        finally:
            dA.free()
            dB.free()
    finally:
        for k, v in saved.items():
            bm.set_param(k, v)
