// Prints the strip schedule's launch shape (mcrg::strip_shape) for each line
// "L nl n_wg n_tail measure halo_ok split_ok" read on stdin, one line of
// "name=value" fields per shape, for tests/test_strip_shape_cpu.py.
#include <cstdio>

#include "../../mcrg_amd/csrc/strip.hpp"

int main()
{
    std::printf("const kStripRing=%u kStripSnapSlots=%u kStripFirstSlots=%u kStripFirstSlotsSplit=%u "
                "kStripCntWords=%u kRedWords=%u kStaticScratchWords=%zu kStripFlagStride=%u\n",
                mcrg::kStripRing, mcrg::kStripSnapSlots, mcrg::kStripFirstSlots, mcrg::kStripFirstSlotsSplit,
                mcrg::kStripCntWords, mcrg::kRedWords, mcrg::kStaticScratchWords, mcrg::kStripFlagStride);
    unsigned L, nl, n_wg, n_tail, measure, halo_ok, split_ok;
    while (std::scanf("%u %u %u %u %u %u %u", &L, &nl, &n_wg, &n_tail, &measure, &halo_ok, &split_ok) == 7) {
        const mcrg::StripShape s = mcrg::strip_shape(L, nl, n_wg, n_tail, measure != 0, halo_ok != 0, split_ok != 0);
        std::printf("shape L=%u nl=%u n_wg=%u R=%u wpr=%u lam=%u n_tail=%u measure=%u threads=%u overlap=%u halo=%u "
                    "split=%u n_meas=%u snap_words=%u first_slots=%u o_err=%u o_consumed=%u o_xhalo=%u "
                    "o_xfirst=%u o_ring=%u o_snapc=%u o_snap=%u ws_words=%u cnt_words=%u first_words=%u "
                    "lam_words=%u per_wg=%u lds_strip=%u lds_tail=%u lds_words=%u",
                    s.L, s.nl, s.n_wg, s.R, s.wpr, s.lam, s.n_tail, s.measure, s.threads, s.overlap, s.halo,
                    s.split, s.n_meas, s.snap_words, s.first_slots, s.o_err, s.o_consumed, s.o_xhalo, s.o_xfirst,
                    s.o_ring, s.o_snapc, s.o_snap, s.ws_words, s.cnt_words, s.first_words, s.lam_words, s.per_wg,
                    s.lds_strip, s.lds_tail, s.lds_words);
        for (unsigned l = 0; l <= mcrg::kMaxLevels; ++l) std::printf(" row_off%u=%u", l, s.row_off[l]);
        std::printf("\n");
    }
    return 0;
}
