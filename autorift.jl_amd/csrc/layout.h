// LDS and scratch layouts of the tracker kernels, shared by the kernels and the host dispatcher.
#pragma once
#include <stddef.h>
#include <stdint.h>


namespace ar {

struct U8Layout {
    int ldw, rows, cols;   // window: column pitch ldw bytes (int8, value ^ 0x80), column-major
    int cpad;              // chip: cols_t = max_cw + 2 columns of cpad bytes; row u at 16 + u
    int cols_t;
    bool acc_in_w;         // shift accumulators reuse the window's bytes once the MFMAs are done
    int acc_cols_t;        // chip columns the accumulators may cover past the window (then their
                           // zero rows are restored before the next chip is staged)
    size_t off_w, off_t, off_s1, off_s2, off_acc, off_tmp, total;
};

__host__ __device__ inline int rup(int x, int m) { return (x + m - 1) / m * m; }

// Matrix-core batches of 4 row-shift tiles per wave: with at most one per wave (and the templated
// band loop), a wave holds its accumulators in registers until every wave is past the window.
__host__ __device__ inline bool u8_one_batch(int max_ch, int max_rx, int max_ry) {
    int nti = (2 * max_ry + 15) / 16, ntj = (2 * max_rx + 15) / 16;
    int nrb = (max_ch + 15 + 15) / 16;
    return nrb <= 9 && ((nti + 3) / 4) * ntj <= 4;
}

// LDS of the u8 tracker. The matrix-core loop reads whole 16x16 tiles of the window: rows up to
// 16 (nti + nrb) and columns up to 16 ntj + max_cw + 1. Only the max_cw + 2 max_rx - 1 real
// columns are allocated; reads past them land on the chip and the sums, finite int8 data that
// meets a zero chip column or a dropped shift. Chip fragments are 16-byte reads at rows
// 16 rb - l16 + 16, rb < nrb, so a chip column needs 16 nrb + 16 bytes.
// Regions: window, chip, S1, S2 (int32 per shift), then the per-shift int32 accumulators -- in
// the window's bytes (running on into the chip's) when they fit there and a wave holds one batch
// -- and, when the pass builds its box sums per point (box == false), the column sums of W / W^2.
// sglobal: the box sums S1 / S2 live in a global scratch slot per workgroup instead of LDS (the
// batched tracker's largest radius class, whose LDS would otherwise admit one workgroup per CU).
__host__ __device__ inline U8Layout u8_layout(int max_ch, int max_cw, int max_rx, int max_ry,
                                             bool box, bool sglobal = false) {
    U8Layout L;
    int nti = (2 * max_ry + 15) / 16, ntj = (2 * max_rx + 15) / 16;
    int nrb = (max_ch + 15 + 15) / 16;   // 16-row blocks of the band, ceil((ch + 15) / 16)
    L.rows = 16 * (nti + nrb);
    // an odd number of 16-byte units per column keeps the 16 lanes of a b128 read on
    // distinct LDS banks
    L.ldw = ((L.rows / 16) & 1) ? L.rows : L.rows + 16;
    L.cols = 16 * ntj + max_cw + 2;   // columns the tile loop reads
    L.cpad = 16 * nrb + 16;
    L.cols_t = max_cw + 2;
    int nr = 2 * max_ry, nc = 2 * max_rx;
    int winw = max_cw + 2 * max_rx - 1;
    size_t shifts = (size_t)nr * nc * 4;
    L.off_w = 0;
    L.off_t = (size_t)winw * L.ldw;
    L.off_s1 = rup((int)(L.off_t + (size_t)(max_cw + 2) * L.cpad), 16);
    L.off_s2 = L.off_s1 + (sglobal ? 0 : shifts);
    L.off_tmp = L.off_s2 + (sglobal ? 0 : shifts);
    L.acc_in_w = shifts <= L.off_s1 && u8_one_batch(max_ch, max_rx, max_ry);
    L.acc_cols_t = 0;
    if (L.acc_in_w && shifts > L.off_t) L.acc_cols_t = (int)((shifts - L.off_t + L.cpad - 1) / L.cpad);
    size_t colsum = box ? 0 : (size_t)nr * winw * 4;   // column sums of W, then of W^2
    size_t tmp = colsum;
    if (L.acc_in_w) {
        L.off_acc = L.off_w;
    } else {
        L.off_acc = L.off_tmp;                          // the column sums are dead by then
        if (shifts > tmp) tmp = shifts;
    }
    L.total = L.off_tmp + tmp;
    size_t reach = (size_t)L.cols * L.ldw + 64;   // past the furthest window byte the tile loop reads
    if (reach > L.total) L.total = reach;
    return L;
}

struct GenScratch {
    size_t off_chip, off_win, off_s, off_s2, off_surf, off_chip0, off_best, total;
};

// Per-workgroup global scratch of the generic tracker: the chip being correlated, the window,
// the two summed-area tables, the surface, and for a rotation search the unrotated chip and the
// best surface so far.
__host__ __device__ inline GenScratch gen_scratch(int max_ch, int max_cw, int max_rx, int max_ry) {
    GenScratch G;
    int winh = max_ch + 2 * max_ry - 1, winw = max_cw + 2 * max_rx - 1;
    G.off_chip = 0;
    G.off_win = rup(max_ch * max_cw * 4, 256);
    G.off_s = G.off_win + rup(winh * winw * 4, 256);
    G.off_s2 = G.off_s + rup((winh + 1) * (winw + 1) * 8, 256);
    G.off_surf = G.off_s2 + rup((winh + 1) * (winw + 1) * 8, 256);
    G.off_chip0 = G.off_surf + rup(4 * max_rx * max_ry * 4, 256);
    G.off_best = G.off_chip0 + rup(max_ch * max_cw * 4, 256);
    G.total = G.off_best + rup(4 * max_rx * max_ry * 4, 256);
    return G;
}

// ---- mixed-precision tracker (BF16 / FP16 operands, FP32 accumulation)

// Band heights (in 8-row blocks, ceil((ch + 15) / 8) rounded up) with a dedicated MFMA loop;
// chips up to 128 rows. 0: the chip is too tall for the mixed-precision kernel.
__host__ __device__ inline int mp_nrb(int ch) {
    const int need = (ch + 15 + 7) / 8;
    const int opts[6] = {4, 6, 8, 10, 14, 18};
    for (int k = 0; k < 6; k++)
        if (need <= opts[k]) return opts[k];
    return 0;
}

struct MpLayout {
    int nti, ntj, nbt;   // row-shift tiles, column-shift tiles, batches of 4 row-shift tiles
    int ldw;             // window column pitch in 16-bit elements (an odd multiple of 8)
    int cols;            // window columns stored: every column the tile loop reads
    int cpad;            // chip column pitch in elements; chip row u at 16 + u, zeros around it
    int cw4;             // chip columns stored (a multiple of 4; the extra ones zero)
    bool num_in_w;       // numerators written over the window once every wave is past the
                         // MFMAs (each wave holds one tile batch in registers until then)
    size_t off_w, off_num, off_c, off_pc, total;   // byte offsets
};

// LDS of track_mp_kernel: window (16-bit, column-major, rows past the window zero), the shift
// numerators (FP32, column-major over the row shift i) and the 16-bit chip. The Float32 chip is
// staged first in the window's bytes; after the argmax the LDS from offset 0 holds the Float32
// window block under the 5x5 patch ((ch + 4) x (cw + 4)) and the mean-removed chip (at off_pc)
// for the exact Float64 values.
__host__ __device__ inline MpLayout mp_layout(int max_ch, int max_cw, int max_rx, int max_ry,
                                             int nrb) {
    MpLayout L;
    L.nti = (2 * max_ry + 15) / 16;
    L.ntj = (2 * max_rx + 15) / 16;
    L.nbt = (L.nti + 3) / 4;
    const int rows = 64 * L.nbt + 8 * nrb;
    L.ldw = ((rows / 8) & 1) ? rows : rows + 8;
    L.cw4 = (max_cw + 3) / 4 * 4;
    L.cols = 16 * L.ntj + L.cw4;
    L.cpad = 8 * nrb + 24;
    const size_t wbytes = (size_t)L.cols * L.ldw * 2;
    const size_t numbytes = (size_t)(64 * L.nbt) * (16 * L.ntj) * 4;
    const size_t chipf = (size_t)max_ch * max_cw * 4;
    L.off_w = 0;
    L.num_in_w = L.ntj * L.nbt <= 4 && numbytes <= wbytes;
    L.off_num = L.num_in_w ? 0 : (size_t)rup((int)wbytes, 16);
    L.off_c = (size_t)rup((int)(L.num_in_w ? wbytes : L.off_num + numbytes), 16);
    L.off_pc = (size_t)rup((max_ch + 4) * (max_cw + 4) * 4, 16);
    const size_t post = L.off_pc + chipf;
    if (L.off_c < chipf) L.off_c = (size_t)rup((int)chipf, 16);   // raw chip read while Cl is built
    L.total = L.off_c + (size_t)L.cw4 * L.cpad * 2;
    if (L.total < post) L.total = post;   // the post-argmax region may cover the 16-bit chip
    return L;
}

// ---- direct Float64 tracker (track_f64.hip)

// Slot of one point in the staging buffer (byte offsets; sizes from the pass maxima).
struct F64Slot {
    size_t off_chip, off_b1, off_b2, bytes;
    int chp;   // chip rows rounded up to a multiple of 8 (rows past ch are zero)
};

__host__ __device__ inline F64Slot f64_slot(int max_ch, int max_cw, int max_rx, int max_ry) {
    F64Slot s;
    s.chp = (max_ch + 7) / 8 * 8;
    const size_t shifts = (size_t)(2 * max_rx) * (2 * max_ry);
    s.off_chip = 64;   // header: status (int32), then cnorm (double) at 8
    s.off_b1 = s.off_chip + (size_t)s.chp * max_cw * 8;
    s.off_b2 = s.off_b1 + shifts * 8;
    s.bytes = (s.off_b2 + shifts * 8 + 255) / 256 * 256;
    return s;
}

// LDS column pitch of the window in track_f64_kernel (floats): row r of a column at r + 1, so
// that the 8 rows a lane reads at a time start on a 16-byte boundary; every row a tile reads
// (8 ceil(NR / 8) + chp - 1 of them) lies inside the column, and the pitch is 4 mod 8 floats so
// that neighbouring columns start on different LDS banks.
__host__ __device__ inline int f64_pitch(int max_ch, int max_ry) {
    const int nrp = (2 * max_ry + 7) / 8 * 8, chp = (max_ch + 7) / 8 * 8;
    int p = nrp + chp + 8;
    p = (p + 3) / 4 * 4;
    if (p % 8 == 0) p += 4;
    return p;
}

#define F64_PREP_WG 256

struct F64Stage {
    unsigned char* slots;      // chunk's slots, slot.bytes apart
    size_t slot_bytes;
    size_t off_chip, off_b1, off_b2;
    int chp;
    int64_t p0, cnt;           // chunk: points p0 .. p0 + cnt - 1
    double* tables;            // per-block summed-area table (when it does not fit in LDS)
    size_t table_doubles;      // doubles per table ((winh + 1) x (winw + 1) of the maxima)
};

}  // namespace ar
