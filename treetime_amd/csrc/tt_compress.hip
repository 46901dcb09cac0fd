// Pattern compression of a dense alignment on the device.
//
// Rules (reference treetime/sequence_data.py:325-464, dense branch, and the
// overhang fill of seq_utils.py:196-202):
//   * optionally, leading and trailing '-' of every sequence become the
//     ambiguous character (a sequence of gaps only becomes all ambiguous);
//   * a column whose characters other than the ambiguous one are all the same
//     letter is constant in that letter (its ambiguous characters are
//     rewritten, sequence_data.py:387-394); a column of ambiguous characters
//     only is constant in the ambiguous character;
//   * every other column is variable and gets a pattern of its own
//     (sequence_data.py:396-402);
//   * constant columns of the same letter share one pattern; patterns are
//     numbered in order of their first column; multiplicity = number of
//     columns of the pattern (sequence_data.py:445-447).
//
// Layout: the alignment is sequence-major, n_rows rows of `pitch` bytes
// (column c of row r at r * pitch + c).  Passes:
//   k_row_ends     first / last non-gap column of each row (fill only; scans
//                  from both ends and stops at the first non-gap byte)
//   k_col_stats    per row slice, min and max of each column's non-ambiguous
//                  bytes (16 columns per thread, one 16-byte load per row);
//                  optionally a 256-bin histogram of all bytes
//   k_col_key      reduce the slices: key = letter of a constant column or -1;
//                  first column and column count of every letter
//   k_open_count / k_scan_blocks / k_assign
//                  pattern number of each column that opens a pattern (a
//                  variable column, or the first column of a letter) by an
//                  exclusive scan over 1024-column blocks
//   k_finish       pattern of the other constant columns; multiplicities
//   k_gather       compressed rows [n_rows][P]: the letter for constant
//                  patterns, the column's bytes for variable ones
// Every result is an exact integer function of the input: no order of
// evaluation changes any output.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "tt_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kScanCols = 4 * kThreads;  // columns per block of the scan kernels
constexpr int kGap = '-';

__device__ inline int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// Inclusive scan over the block's 256 threads; `wsum` holds 4 ints of LDS.
// Every thread of the block must call it.
__device__ inline int block_inclusive_scan(int v, int *wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = wave_inclusive_scan(v);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    for (int k = 0; k < w; ++k) incl += wsum[k];
    __syncthreads();
    return incl;
}

__global__ __launch_bounds__(kThreads) void k_init(int *first, int *count, unsigned long long *hist) {
    const int t = threadIdx.x;
    first[t] = INT_MAX;
    count[t] = 0;
    hist[t] = 0ull;
}

// One wave per row: ends[r] = (first non-gap column, last non-gap column),
// (L, -1) for a row of gaps only.
__global__ __launch_bounds__(kThreads) void k_row_ends(const uint8_t *__restrict__ A, int64_t pitch, int n_rows, int L,
                                                      int2 *__restrict__ ends) {
    const int r = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n_rows) return;
    const uint8_t *row = A + (int64_t)r * pitch;
    int first = L, last = -1;
    for (int c0 = 0; c0 < L; c0 += 64) {
        const int c = c0 + lane;
        const unsigned long long m = __ballot(c < L && row[c] != kGap);
        if (m) {
            first = c0 + __ffsll((long long)m) - 1;
            break;
        }
    }
    if (first < L) {
        for (int e = L; e > 0; e -= 64) {
            const int c = e - 1 - lane;
            const unsigned long long m = __ballot(c >= 0 && row[c] != kGap);
            if (m) {
                last = e - 1 - (__ffsll((long long)m) - 1);
                break;
            }
        }
    }
    if (lane == 0) ends[r] = make_int2(first, last);
}

template <bool FILL>
__device__ inline uint32_t filled(uint32_t b, int c, int2 e, int amb) {
    if (FILL && (c < e.x || c > e.y)) return (uint32_t)amb;
    return b;
}

template <int V, bool FILL, bool HIST>
__device__ inline void stats_row(const uint8_t *__restrict__ row, int c0, int L, int2 e, int amb, uint32_t *lo,
                                 uint32_t *hi, unsigned int *h) {
    if constexpr (V == 16) {
        const uint4 w = *reinterpret_cast<const uint4 *>(row + c0);
        const uint32_t word[4] = {w.x, w.y, w.z, w.w};
        const bool inside = !FILL || (c0 >= e.x && c0 + 15 <= e.y);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            uint32_t b = (word[k >> 2] >> (8 * (k & 3))) & 0xffu;
            if (!inside) b = filled<FILL>(b, c0 + k, e, amb);
            const bool a = (int)b == amb;
            lo[k] = min(lo[k], a ? 255u : b);
            hi[k] = max(hi[k], a ? 0u : b);
            if (HIST && c0 + k < L) atomicAdd(&h[b], 1u);
        }
    } else {
        const uint32_t b = filled<FILL>(row[c0], c0, e, amb);
        const bool a = (int)b == amb;
        lo[0] = min(lo[0], a ? 255u : b);
        hi[0] = max(hi[0], a ? 0u : b);
        if (HIST) atomicAdd(&h[b], 1u);
    }
}

// part[slice][c] = lo | hi << 8 over the slice's rows; lo > hi when the
// slice holds only ambiguous bytes in column c.  V = 16 needs pitch % 16 == 0
// and a 16-byte aligned base (columns up to the next multiple of 16 are read
// and their results discarded); V = 1 reads one byte per thread and row.
template <int V, bool FILL, bool HIST>
__global__ __launch_bounds__(kThreads) void k_col_stats(const uint8_t *__restrict__ A, int64_t pitch, int n_rows,
                                                       int L, int rows_per_slice, int amb,
                                                       const int2 *__restrict__ ends, uint16_t *__restrict__ part,
                                                       int part_pitch, unsigned long long *__restrict__ hist) {
    __shared__ unsigned int h[HIST ? 256 : 1];
    if (HIST) {
        h[threadIdx.x] = 0u;
        __syncthreads();
    }
    const int c0 = (blockIdx.x * kThreads + threadIdx.x) * V;
    if (c0 < L) {
        const int r0 = blockIdx.y * rows_per_slice;
        const int r1 = min(n_rows, r0 + rows_per_slice);
        uint32_t lo[V], hi[V];
#pragma unroll
        for (int k = 0; k < V; ++k) {
            lo[k] = 255u;
            hi[k] = 0u;
        }
        const int2 none = make_int2(0, INT_MAX);
        int r = r0;
        for (; r + 4 <= r1; r += 4) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                stats_row<V, FILL, HIST>(A + (int64_t)(r + q) * pitch, c0, L, FILL ? ends[r + q] : none, amb, lo, hi,
                                         h);
        }
        for (; r < r1; ++r) stats_row<V, FILL, HIST>(A + (int64_t)r * pitch, c0, L, FILL ? ends[r] : none, amb, lo, hi, h);
        uint16_t *out = part + (int64_t)blockIdx.y * part_pitch + c0;
        if constexpr (V == 16) {
            // part_pitch is a multiple of 16, so the 32 bytes stay inside the row
            uint32_t pk[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                pk[k] = (lo[2 * k] | hi[2 * k] << 8) | (lo[2 * k + 1] | hi[2 * k + 1] << 8) << 16;
            reinterpret_cast<uint4 *>(out)[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
            reinterpret_cast<uint4 *>(out)[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
        } else {
            out[0] = (uint16_t)(lo[0] | hi[0] << 8);
        }
    }
    if (HIST) {
        __syncthreads();
        if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
    }
}

__global__ __launch_bounds__(kThreads) void k_col_key(const uint16_t *__restrict__ part, int part_pitch, int S, int L,
                                                     int amb, int16_t *__restrict__ key, int *__restrict__ first,
                                                     int *__restrict__ count) {
    __shared__ int s_first[256], s_count[256];
    s_first[threadIdx.x] = INT_MAX;
    s_count[threadIdx.x] = 0;
    __syncthreads();
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j < L) {
        uint32_t lo = 255u, hi = 0u;
        for (int s = 0; s < S; ++s) {
            const uint32_t v = part[(int64_t)s * part_pitch + j];
            lo = min(lo, v & 0xffu);
            hi = max(hi, v >> 8);
        }
        // lo > hi: only ambiguous bytes (amb >= 0 then, since every byte counts otherwise)
        const int k = lo > hi ? amb : (lo == hi ? (int)lo : -1);
        key[j] = (int16_t)k;
        if (k >= 0) {
            atomicMin(&s_first[k], j);
            atomicAdd(&s_count[k], 1);
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (s_count[t]) {
        atomicMin(&first[t], s_first[t]);
        atomicAdd(&count[t], s_count[t]);
    }
}

__device__ inline bool opens(const int16_t *key, const int *first, int j) {
    const int k = key[j];
    return k < 0 || first[k] == j;
}

__global__ __launch_bounds__(kThreads) void k_open_count(const int16_t *__restrict__ key,
                                                        const int *__restrict__ first, int L,
                                                        int *__restrict__ block_open) {
    __shared__ int wsum[4];
    const int base = blockIdx.x * kScanCols + threadIdx.x * 4;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (base + q < L) cnt += opens(key, first, base + q);
    const int incl = block_inclusive_scan(cnt, wsum);
    if (threadIdx.x == kThreads - 1) block_open[blockIdx.x] = incl;
}

// One block: exclusive scan of the per-block counts, total pattern count.
__global__ __launch_bounds__(kThreads) void k_scan_blocks(const int *__restrict__ block_open, int nb,
                                                         int *__restrict__ block_base, int *__restrict__ total) {
    __shared__ int wsum[4];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += kThreads) {
        const int b = b0 + threadIdx.x;
        const int v = b < nb ? block_open[b] : 0;
        const int incl = block_inclusive_scan(v, wsum);
        const int c = carry;
        if (b < nb) block_base[b] = c + incl - v;
        __syncthreads();
        if (threadIdx.x == kThreads - 1) carry = c + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kThreads) void k_assign(const int16_t *__restrict__ key, const int *__restrict__ first,
                                                    const int *__restrict__ block_base, int L, int *__restrict__ pat,
                                                    int *__restrict__ col_of_pat, int *__restrict__ letter_pat) {
    __shared__ int wsum[4];
    const int base = blockIdx.x * kScanCols + threadIdx.x * 4;
    bool op[4];
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        op[q] = base + q < L && opens(key, first, base + q);
        cnt += op[q];
    }
    const int incl = block_inclusive_scan(cnt, wsum);
    int rank = block_base[blockIdx.x] + incl - cnt;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (op[q]) {
            const int j = base + q;
            pat[j] = rank;
            col_of_pat[rank] = j;
            const int k = key[j];
            if (k >= 0) letter_pat[k] = rank;
            ++rank;
        }
}

__global__ __launch_bounds__(kThreads) void k_finish(const int16_t *__restrict__ key, const int *__restrict__ first,
                                                    const int *__restrict__ letter_pat, const int *__restrict__ count,
                                                    const int *__restrict__ col_of_pat, const int *__restrict__ total,
                                                    int L, int *__restrict__ pat, double *__restrict__ mult) {
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= L) return;
    const int k = key[j];
    if (k >= 0 && first[k] != j) pat[j] = letter_pat[k];
    if (j < *total) {
        const int kc = key[col_of_pat[j]];
        mult[j] = kc < 0 ? 1.0 : (double)count[kc];
    }
}

// out[r][p] for the rows of one slice; one thread per pattern.
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_gather(const uint8_t *__restrict__ A, int64_t pitch, int n_rows,
                                                    const int *__restrict__ col_of_pat,
                                                    const int16_t *__restrict__ key, int P, int rows_per_slice,
                                                    const int2 *__restrict__ ends, int amb,
                                                    uint8_t *__restrict__ out) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= P) return;
    const int c = col_of_pat[p];
    const int k = key[c];
    const int r0 = blockIdx.y * rows_per_slice;
    const int r1 = min(n_rows, r0 + rows_per_slice);
    if (k >= 0) {
        for (int r = r0; r < r1; ++r) out[(int64_t)r * P + p] = (uint8_t)k;
    } else {
        for (int r = r0; r < r1; ++r) {
            uint32_t b = A[(int64_t)r * pitch + c];
            if (FILL) b = filled<true>(b, c, ends[r], amb);
            out[(int64_t)r * P + p] = (uint8_t)b;
        }
    }
}

int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Row slices so that a launch has about `target` threads of `per_slice`
// threads each, with at least `min_rows` rows per slice.
int pick_rows_per_slice(int n_rows, int64_t per_slice, int64_t target, int min_rows) {
    int64_t S = std::max<int64_t>(1, (target + per_slice - 1) / per_slice);
    S = std::min<int64_t>(S, std::max(1, n_rows / std::max(1, min_rows)));
    return ceil_div(n_rows, S);
}

}  // namespace

namespace ttk {

CompressLayout compress_layout(int n_rows, int L) {
    CompressLayout l{};
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int V = 16;
    const int per_slice = ceil_div(L, V);
    l.rows_per_slice = pick_rows_per_slice(n_rows, per_slice, 131072, 32);
    l.slices = ceil_div(n_rows, l.rows_per_slice);
    l.part_pitch = ((L + 15) / 16) * 16;
    l.n_blocks = ceil_div(L, kScanCols);
    size_t o = 0;
    l.off_ends = o;
    o += up((size_t)n_rows * sizeof(int2));
    l.off_part = o;
    o += up((size_t)l.slices * l.part_pitch * sizeof(uint16_t));
    l.off_key = o;
    o += up((size_t)L * sizeof(int16_t));
    l.off_small = o;  // first, count, letter_pat (256 ints each), total
    o += up(3 * 256 * sizeof(int) + sizeof(int));
    l.off_hist = o;
    o += up(256 * sizeof(unsigned long long));
    l.off_blocks = o;
    o += up((size_t)2 * l.n_blocks * sizeof(int));
    l.off_col_of_pat = o;
    o += up((size_t)L * sizeof(int));
    l.total = o;
    return l;
}

hipError_t compress_columns(const uint8_t *A, int64_t pitch, int n_rows, int L, int amb, bool fill, bool hist,
                            void *scratch, const CompressLayout &l, int *pat, hipStream_t s) {
    char *base = static_cast<char *>(scratch);
    int2 *ends = reinterpret_cast<int2 *>(base + l.off_ends);
    uint16_t *part = reinterpret_cast<uint16_t *>(base + l.off_part);
    int16_t *key = reinterpret_cast<int16_t *>(base + l.off_key);
    int *first = reinterpret_cast<int *>(base + l.off_small);
    int *count = first + 256;
    int *letter_pat = count + 256;
    int *total = letter_pat + 256;
    unsigned long long *hst = reinterpret_cast<unsigned long long *>(base + l.off_hist);
    int *block_open = reinterpret_cast<int *>(base + l.off_blocks);
    int *block_base = block_open + l.n_blocks;
    int *col_of_pat = reinterpret_cast<int *>(base + l.off_col_of_pat);

    k_init<<<1, kThreads, 0, s>>>(first, count, hst);
    if (fill) k_row_ends<<<ceil_div(n_rows, kThreads / 64), kThreads, 0, s>>>(A, pitch, n_rows, L, ends);
    const bool v16 = pitch % 16 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    const int rps = v16 ? l.rows_per_slice : pick_rows_per_slice(n_rows, L, 131072, 32);
    const int S = ceil_div(n_rows, rps);
    if (S > l.slices) return hipErrorInvalidValue;  // layout sized for the 16-byte path's slice count or more
    const dim3 grid(ceil_div(L, (v16 ? 16 : 1) * kThreads), S);
#define TT_STATS(V_, F_, H_) \
    k_col_stats<V_, F_, H_><<<grid, kThreads, 0, s>>>(A, pitch, n_rows, L, rps, amb, ends, part, l.part_pitch, hst)
    if (v16) {
        if (fill) {
            if (hist) TT_STATS(16, true, true); else TT_STATS(16, true, false);
        } else {
            if (hist) TT_STATS(16, false, true); else TT_STATS(16, false, false);
        }
    } else {
        if (fill) {
            if (hist) TT_STATS(1, true, true); else TT_STATS(1, true, false);
        } else {
            if (hist) TT_STATS(1, false, true); else TT_STATS(1, false, false);
        }
    }
#undef TT_STATS
    k_col_key<<<ceil_div(L, kThreads), kThreads, 0, s>>>(part, l.part_pitch, S, L, amb, key, first, count);
    k_open_count<<<l.n_blocks, kThreads, 0, s>>>(key, first, L, block_open);
    k_scan_blocks<<<1, kThreads, 0, s>>>(block_open, l.n_blocks, block_base, total);
    k_assign<<<l.n_blocks, kThreads, 0, s>>>(key, first, block_base, L, pat, col_of_pat, letter_pat);
    return hipGetLastError();
}

hipError_t compress_finish(const uint8_t *A, int64_t pitch, int n_rows, int L, int amb, bool fill, void *scratch,
                           const CompressLayout &l, int P, int *pat, double *mult, uint8_t *out, hipStream_t s) {
    char *base = static_cast<char *>(scratch);
    const int2 *ends = reinterpret_cast<const int2 *>(base + l.off_ends);
    const int16_t *key = reinterpret_cast<const int16_t *>(base + l.off_key);
    const int *first = reinterpret_cast<const int *>(base + l.off_small);
    const int *count = first + 256;
    const int *letter_pat = count + 256;
    const int *total = letter_pat + 256;
    const int *col_of_pat = reinterpret_cast<const int *>(base + l.off_col_of_pat);
    k_finish<<<ceil_div(L, kThreads), kThreads, 0, s>>>(key, first, letter_pat, count, col_of_pat, total, L, pat, mult);
    if (P > 0) {
        const int rps = pick_rows_per_slice(n_rows, P, 131072, 8);
        const dim3 grid(ceil_div(P, kThreads), ceil_div(n_rows, rps));
        if (fill)
            k_gather<true><<<grid, kThreads, 0, s>>>(A, pitch, n_rows, col_of_pat, key, P, rps, ends, amb, out);
        else
            k_gather<false><<<grid, kThreads, 0, s>>>(A, pitch, n_rows, col_of_pat, key, P, rps, ends, amb, out);
    }
    return hipGetLastError();
}

const int *compress_total_ptr(void *scratch, const CompressLayout &l) {
    return reinterpret_cast<const int *>(static_cast<char *>(scratch) + l.off_small) + 3 * 256;
}

const unsigned long long *compress_hist_ptr(void *scratch, const CompressLayout &l) {
    return reinterpret_cast<const unsigned long long *>(static_cast<char *>(scratch) + l.off_hist);
}

}  // namespace ttk
