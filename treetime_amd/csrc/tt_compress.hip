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
//   k_prologue     clears the tables; with the fill, first / last non-gap
//                  column of each row (scans from both ends and stops at the
//                  first non-gap byte)
//   k_col_stats    per row slice, min and max of each column's non-ambiguous
//                  bytes (16 columns per thread, one 16-byte load per row,
//                  packed 16-bit min / max); optionally a 256-bin histogram
//                  of all bytes
//   k_col_key      reduce the slices: key = letter of a constant column or -1;
//                  first column and column count of every letter
//   k_open_scan / k_assign
//                  pattern number of each column that opens a pattern (a
//                  variable column, or the first column of a letter) by an
//                  exclusive scan over 1024-column blocks (the last block of
//                  k_open_scan scans the block counts)
//   k_gather       pattern of the other constant columns, multiplicities, and
//                  the compressed rows [n_rows][P]: the letter for constant
//                  patterns, the column's bytes for variable ones
// Every result is an exact integer function of the input: no order of
// evaluation changes any output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdlib>

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

// Block 0 clears the per-letter tables, the byte histogram and the block
// counter; with FILL, one wave per row also finds ends[r] = (first non-gap
// column, last non-gap column), (L, -1) for a row of gaps only.
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_prologue(const uint8_t *__restrict__ A, int64_t pitch, int n_rows, int L,
                                                      int *__restrict__ first, int *__restrict__ count,
                                                      unsigned long long *__restrict__ hist,
                                                      int *__restrict__ counter, int2 *__restrict__ ends) {
    if (blockIdx.x == 0) {
        const int t = threadIdx.x;
        first[t] = INT_MAX;
        count[t] = 0;
        hist[t] = 0ull;
        if (t == 0) *counter = 0;
    }
    if (!FILL) return;
    const int r = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n_rows) return;
    const uint8_t *row = A + (int64_t)r * pitch;
    int first_c = L, last_c = -1;
    for (int c0 = 0; c0 < L; c0 += 64) {
        const int c = c0 + lane;
        const unsigned long long m = __ballot(c < L && row[c] != kGap);
        if (m) {
            first_c = c0 + __ffsll((long long)m) - 1;
            break;
        }
    }
    if (first_c < L) {
        for (int e = L; e > 0; e -= 64) {
            const int c = e - 1 - lane;
            const unsigned long long m = __ballot(c >= 0 && row[c] != kGap);
            if (m) {
                last_c = e - 1 - (__ffsll((long long)m) - 1);
                break;
            }
        }
    }
    if (lane == 0) ends[r] = make_int2(first_c, last_c);
}

template <bool FILL>
__device__ inline uint32_t filled(uint32_t b, int c, int2 e, int amb) {
    if (FILL && (c < e.x || c > e.y)) return (uint32_t)amb;
    return b;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ inline u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }

// Column statistics of 16 consecutive columns, two columns per 16-bit lane
// pair: lo[2q] / hi[2q] hold the even bytes (columns 4q, 4q + 2) of word q,
// lo[2q + 1] / hi[2q + 1] the odd bytes (4q + 1, 4q + 3).
//   MASKED (an ambiguous character a exists): x = byte ^ a, so ambiguous
//   bytes are 0; hi = max x (0 if all ambiguous), lo = min (x - 1) in 16 bits
//   (an ambiguous byte gives 0xffff and never wins).  The column is constant
//   iff lo + 1 == hi, in letter hi ^ a; all ambiguous iff hi == 0.
//   raw (no ambiguous character): lo = min byte, hi = max byte.
template <bool MASKED>
struct Acc16 {
    u16x2 lo[8], hi[8];
    __device__ void init() {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            lo[i] = (u16x2)0xffff;
            hi[i] = (u16x2)0;
        }
    }
    __device__ void add(const uint32_t (&x)[4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const u16x2 e = as_u16x2(x[q] & 0x00ff00ffu);
            const u16x2 o = as_u16x2((x[q] >> 8) & 0x00ff00ffu);
            if (MASKED) {
                lo[2 * q] = __builtin_elementwise_min(lo[2 * q], e - (u16x2)1);
                lo[2 * q + 1] = __builtin_elementwise_min(lo[2 * q + 1], o - (u16x2)1);
            } else {
                lo[2 * q] = __builtin_elementwise_min(lo[2 * q], e);
                lo[2 * q + 1] = __builtin_elementwise_min(lo[2 * q + 1], o);
            }
            hi[2 * q] = __builtin_elementwise_max(hi[2 * q], e);
            hi[2 * q + 1] = __builtin_elementwise_max(hi[2 * q + 1], o);
        }
    }
    // packed partial of column 4q + b: low byte lo, high byte hi
    __device__ uint32_t packed(int q, int b) const {
        const u16x2 l = lo[2 * q + (b & 1)], h = hi[2 * q + (b & 1)];
        const uint32_t lv = (b & 2) ? l.y : l.x, hv = (b & 2) ? h.y : h.x;
        return (lv & 0xffu) | hv << 8;
    }
};

// Bytes of word q (columns c .. c + 3) inside the row's non-overhang range
// [e.x, e.y] keep their value; the others become 0 (ambiguous after the xor).
__device__ inline uint32_t keep_mask(int c, int2 e) {
    uint32_t m = 0u;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (c + i >= e.x && c + i <= e.y) m |= 0xffu << (8 * i);
    return m;
}

template <bool MASKED, bool FILL, bool HIST>
__device__ inline void stats_row16(const uint4 w, int c0, int L, int2 e, int amb, Acc16<MASKED> &acc,
                                   unsigned int *h) {
    const uint32_t a4 = MASKED ? (uint32_t)amb * 0x01010101u : 0u;
    uint32_t x[4] = {w.x ^ a4, w.y ^ a4, w.z ^ a4, w.w ^ a4};
    if (FILL) {
        const bool inside = c0 >= e.x && c0 + 15 <= e.y;
        // wave-uniform test first: rows without overhangs skip the masking
        if (__ballot(!inside) && !inside) {
            if (c0 > e.y || c0 + 15 < e.x) {  // the whole group is overhang: ambiguous only
#pragma unroll
                for (int q = 0; q < 4; ++q) x[q] = 0u;
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) x[q] &= keep_mask(c0 + 4 * q, e);
            }
        }
    }
    acc.add(x);
    if (HIST) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (c0 + k < L) atomicAdd(&h[((x[k >> 2] >> (8 * (k & 3))) & 0xffu) ^ (a4 & 0xffu)], 1u);
    }
}

// part[slice][c] = packed statistics of column c over the slice's rows (see
// Acc16).  V = 16 needs pitch % 16 == 0 and a 16-byte aligned base (columns
// up to the next multiple of 16 are read and their results discarded); V = 1
// reads one byte per thread and row.
template <int V, bool MASKED, bool FILL, bool HIST>
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
        const int2 none = make_int2(0, INT_MAX);
        uint16_t *out = part + (int64_t)blockIdx.y * part_pitch + c0;
        if constexpr (V == 16) {
            Acc16<MASKED> acc;
            acc.init();
            const uint8_t *p = A + (int64_t)r0 * pitch + c0;
            int r = r0;
            for (; r + 4 <= r1; r += 4, p += 4 * pitch) {
                uint4 w[4];
                int2 e[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    w[q] = *reinterpret_cast<const uint4 *>(p + q * pitch);
                    e[q] = FILL ? ends[r + q] : none;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) stats_row16<MASKED, FILL, HIST>(w[q], c0, L, e[q], amb, acc, h);
            }
            for (; r < r1; ++r, p += pitch)
                stats_row16<MASKED, FILL, HIST>(*reinterpret_cast<const uint4 *>(p), c0, L, FILL ? ends[r] : none,
                                                amb, acc, h);
            // part_pitch is a multiple of 16, so the 32 bytes stay inside the row
            uint32_t pk[8];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                pk[k] = acc.packed((2 * k) >> 2, (2 * k) & 3) | acc.packed((2 * k + 1) >> 2, (2 * k + 1) & 3) << 16;
            reinterpret_cast<uint4 *>(out)[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
            reinterpret_cast<uint4 *>(out)[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
        } else {
            uint32_t lo = 0xffu, hi = 0u;
            for (int r = r0; r < r1; ++r) {
                const uint32_t b = filled<FILL>(A[(int64_t)r * pitch + c0], c0, FILL ? ends[r] : none, amb);
                if (HIST) atomicAdd(&h[b], 1u);
                if (MASKED) {
                    const uint32_t x = b ^ (uint32_t)amb;
                    lo = min(lo, (x - 1u) & 0xffffu);
                    hi = max(hi, x);
                } else {
                    lo = min(lo, b);
                    hi = max(hi, b);
                }
            }
            out[0] = (uint16_t)((lo & 0xffu) | hi << 8);
        }
    }
    if (HIST) {
        __syncthreads();
        if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
    }
}

// Column key from the slices' statistics: the letter of a constant column,
// -1 for a variable one.  Block = 64 columns x 4 groups of slices; the first
// column and the column count of every letter are gathered per block in LDS.
__global__ __launch_bounds__(kThreads) void k_col_key(const uint16_t *__restrict__ part, int part_pitch, int S, int L,
                                                     int amb, int16_t *__restrict__ key, int *__restrict__ first,
                                                     int *__restrict__ count) {
    __shared__ int s_first[256], s_count[256];
    __shared__ uint32_t s_lo[4][64], s_hi[4][64];
    s_first[threadIdx.x] = INT_MAX;
    s_count[threadIdx.x] = 0;
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    uint32_t lo = 0xffu, hi = 0u;
    if (j < L) {
        int s = g;
        for (; s + 28 < S; s += 32) {  // 8 independent loads in flight
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = part[(int64_t)(s + 4 * u) * part_pitch + j];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                lo = min(lo, v[u] & 0xffu);
                hi = max(hi, v[u] >> 8);
            }
        }
        for (; s < S; s += 4) {
            const uint32_t v = part[(int64_t)s * part_pitch + j];
            lo = min(lo, v & 0xffu);
            hi = max(hi, v >> 8);
        }
    }
    s_lo[g][lane] = lo;
    s_hi[g][lane] = hi;
    __syncthreads();
    if (g == 0 && j < L) {
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            lo = min(lo, s_lo[k][lane]);
            hi = max(hi, s_hi[k][lane]);
        }
        int k;
        if (amb >= 0)
            k = hi == 0u ? amb : (lo + 1u == hi ? (int)(hi ^ (uint32_t)amb) : -1);
        else
            k = lo == hi ? (int)lo : -1;
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

// Opening columns per 1024-column block; the last block to finish (block
// counter) scans the per-block counts: block_base[b], total = P.
__global__ __launch_bounds__(kThreads) void k_open_scan(const int16_t *__restrict__ key,
                                                       const int *__restrict__ first, int L, int nb,
                                                       int *block_open, int *__restrict__ block_base,
                                                       int *__restrict__ total, int *counter) {
    __shared__ int wsum[4];
    __shared__ int carry;
    __shared__ bool last;
    const int base = blockIdx.x * kScanCols + threadIdx.x * 4;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (base + q < L) cnt += opens(key, first, base + q);
    const int incl0 = block_inclusive_scan(cnt, wsum);
    if (threadIdx.x == kThreads - 1) {
        block_open[blockIdx.x] = incl0;
        __threadfence();
        last = atomicAdd(counter, 1) == nb - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const volatile int *bo = block_open;  // written by other blocks: read past the CU's cache
    for (int b0 = 0; b0 < nb; b0 += kThreads) {
        const int b = b0 + threadIdx.x;
        const int v = b < nb ? bo[b] : 0;
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

constexpr int kGatherRows = 16;

// Pattern of the constant columns that do not open a pattern, the
// multiplicities, and out[r][p] (row stride P) over items (slice of
// kGatherRows rows, pattern p); grid-stride.  P is read on the device;
// nothing but the column map is written when P > capacity.
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_gather(const uint8_t *__restrict__ A, int64_t pitch, int n_rows, int L,
                                                    const int *__restrict__ col_of_pat,
                                                    const int16_t *__restrict__ key, const int *__restrict__ first,
                                                    const int *__restrict__ letter_pat,
                                                    const int *__restrict__ count, const int *__restrict__ total,
                                                    int64_t capacity, const int2 *__restrict__ ends, int amb,
                                                    int *__restrict__ pat, double *__restrict__ mult,
                                                    uint8_t *__restrict__ out) {
    const int P = *total;
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    const int64_t tid = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const bool fits = P <= capacity;
    for (int64_t j = tid; j < L; j += stride) {
        const int k = key[j];
        if (k >= 0 && first[k] != j) pat[j] = letter_pat[k];
        if (j < P && fits) {
            const int kc = key[col_of_pat[j]];
            mult[j] = kc < 0 ? 1.0 : (double)count[kc];
        }
    }
    if (P <= 0 || !fits) return;
    const int n_slices = (n_rows + kGatherRows - 1) / kGatherRows;
    const int64_t items = (int64_t)n_slices * P;
    for (int64_t i = tid; i < items; i += stride) {
        const int sl = (int)(i / P);
        const int p = (int)(i - (int64_t)sl * P);
        const int c = col_of_pat[p];
        const int k = key[c];
        const int r0 = sl * kGatherRows;
        const int r1 = min(n_rows, r0 + kGatherRows);
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
}

int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Row slices so that a launch has about `target` threads of `per_slice`
// threads each, with at least `min_rows` rows per slice.
int pick_rows_per_slice(int n_rows, int64_t per_slice, int64_t target, int min_rows) {
    int64_t S = std::max<int64_t>(1, (target + per_slice - 1) / per_slice);
    S = std::min<int64_t>(S, std::max(1, n_rows / std::max(1, min_rows)));
    return ceil_div(n_rows, S);
}

// Threads the column-statistics pass aims for (TT_COMPRESS_THREADS
// overrides): enough 16-byte loads in flight to cover HBM latency, few
// enough row slices that the partial statistics stay small.
int64_t stats_threads() {
    static const int64_t v = [] {
        const char *e = getenv("TT_COMPRESS_THREADS");
        const long long x = e ? atoll(e) : 131072;
        return (int64_t)(x > 0 ? x : 131072);
    }();
    return v;
}

}  // namespace

namespace ttk {

CompressLayout compress_layout(int n_rows, int L) {
    CompressLayout l{};
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int V = 16;
    const int per_slice = ceil_div(L, V);
    l.rows_per_slice = pick_rows_per_slice(n_rows, per_slice, stats_threads(), 32);
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
    o += up((size_t)(2 * l.n_blocks + 1) * sizeof(int));
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

    if (fill && amb < 0) return hipErrorInvalidValue;  // the fill writes the ambiguous character
    int *counter = block_base + l.n_blocks;
    if (fill)
        k_prologue<true><<<ceil_div(n_rows, kThreads / 64), kThreads, 0, s>>>(A, pitch, n_rows, L, first, count, hst,
                                                                             counter, ends);
    else
        k_prologue<false><<<1, kThreads, 0, s>>>(A, pitch, n_rows, L, first, count, hst, counter, ends);
    const bool v16 = pitch % 16 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    const int rps = v16 ? l.rows_per_slice : pick_rows_per_slice(n_rows, L, stats_threads(), 32);
    const int S = ceil_div(n_rows, rps);
    if (S > l.slices) return hipErrorInvalidValue;  // layout sized for the 16-byte path's slice count or more
    const dim3 grid(ceil_div(L, (v16 ? 16 : 1) * kThreads), S);
    const int sel = (v16 ? 8 : 0) | (amb >= 0 ? 4 : 0) | (fill ? 2 : 0) | (hist ? 1 : 0);
#define TT_STATS(V_, M_, F_, H_) \
    k_col_stats<V_, M_, F_, H_><<<grid, kThreads, 0, s>>>(A, pitch, n_rows, L, rps, amb, ends, part, l.part_pitch, hst)
    switch (sel) {
        case 8 | 4 | 2 | 1: TT_STATS(16, true, true, true); break;
        case 8 | 4 | 2: TT_STATS(16, true, true, false); break;
        case 8 | 4 | 1: TT_STATS(16, true, false, true); break;
        case 8 | 4: TT_STATS(16, true, false, false); break;
        case 8 | 1: TT_STATS(16, false, false, true); break;
        case 8: TT_STATS(16, false, false, false); break;
        case 4 | 2 | 1: TT_STATS(1, true, true, true); break;
        case 4 | 2: TT_STATS(1, true, true, false); break;
        case 4 | 1: TT_STATS(1, true, false, true); break;
        case 4: TT_STATS(1, true, false, false); break;
        case 1: TT_STATS(1, false, false, true); break;
        default: TT_STATS(1, false, false, false); break;
    }
#undef TT_STATS
    k_col_key<<<ceil_div(L, 64), kThreads, 0, s>>>(part, l.part_pitch, S, L, amb, key, first, count);
    k_open_scan<<<l.n_blocks, kThreads, 0, s>>>(key, first, L, l.n_blocks, block_open, block_base, total, counter);
    k_assign<<<l.n_blocks, kThreads, 0, s>>>(key, first, block_base, L, pat, col_of_pat, letter_pat);
    return hipGetLastError();
}

hipError_t compress_finish(const uint8_t *A, int64_t pitch, int n_rows, int L, int amb, bool fill, void *scratch,
                           const CompressLayout &l, int64_t capacity, int *pat, double *mult, uint8_t *out,
                           hipStream_t s) {
    char *base = static_cast<char *>(scratch);
    const int2 *ends = reinterpret_cast<const int2 *>(base + l.off_ends);
    const int16_t *key = reinterpret_cast<const int16_t *>(base + l.off_key);
    const int *first = reinterpret_cast<const int *>(base + l.off_small);
    const int *count = first + 256;
    const int *letter_pat = count + 256;
    const int *total = letter_pat + 256;
    const int *col_of_pat = reinterpret_cast<const int *>(base + l.off_col_of_pat);
    // grid for the largest possible pattern count; the kernel reads the real one
    const int64_t p_max = std::max<int64_t>(1, std::min<int64_t>(L, capacity));
    const int64_t items = std::max<int64_t>((int64_t)ceil_div(n_rows, kGatherRows) * p_max, L);
    const int blocks = (int)std::min<int64_t>(4096, (items + kThreads - 1) / kThreads);
    if (fill)
        k_gather<true><<<blocks, kThreads, 0, s>>>(A, pitch, n_rows, L, col_of_pat, key, first, letter_pat, count,
                                                   total, capacity, ends, amb, pat, mult, out);
    else
        k_gather<false><<<blocks, kThreads, 0, s>>>(A, pitch, n_rows, L, col_of_pat, key, first, letter_pat, count,
                                                    total, capacity, ends, amb, pat, mult, out);
    return hipGetLastError();
}

const int *compress_total_ptr(void *scratch, const CompressLayout &l) {
    return reinterpret_cast<const int *>(static_cast<char *>(scratch) + l.off_small) + 3 * 256;
}

const unsigned long long *compress_hist_ptr(void *scratch, const CompressLayout &l) {
    return reinterpret_cast<const unsigned long long *>(static_cast<char *>(scratch) + l.off_hist);
}

}  // namespace ttk
