// sort.hip -- stable counting sort of (key, value) pairs with small keys.
//
// The NUFFT set-up orders visibilities and gridding-tile entries by tile
// (the reference's sort_input, c++/mfista_nufft_lib.cpp:550-641, orders by
// grid cell with std::sort).  Keys are tile indices below 2^end_bit (a few
// thousand), so one counting pass replaces a multi-pass radix sort:
//   1. k_seg_count   one wave per contiguous segment of the input builds the
//                    segment's key histogram in LDS -> counts[seg][key]
//   2. k_key_totals  totals[key] = sum over segments (fixed order)
//   3. k_key_scan    exclusive scan over keys (one workgroup)
//   4. k_seg_offsets offsets[seg][key] = start of key + counts of earlier segments
//   5. k_seg_scatter each wave walks its segment again, 64 pairs at a time; a
//                    pair's rank among equal keys earlier in the batch comes
//                    from ballots over the key bits, so equal keys keep their
//                    input order: the sort is stable and deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace mf {

constexpr int kSortMaxBits = 14;  // LDS holds one int per key: 2^14 keys = 64 KiB

struct SortPlan {
  int n, nkeys, seg, nseg;
};

static SortPlan sort_plan(int n, int end_bit) {
  SortPlan p;
  p.n = n;
  p.nkeys = 1 << end_bit;
  // segments of at least 256 pairs, at most 1024 of them
  p.seg = std::max(256, (n + 1023) / 1024);
  p.seg = (p.seg + 63) / 64 * 64;
  p.nseg = std::max(1, (n + p.seg - 1) / p.seg);
  return p;
}

__global__ void __launch_bounds__(64) k_seg_count(const int* __restrict__ keys, int n, int seg, int nkeys,
                                                  int* __restrict__ counts) {
  extern __shared__ int hist[];
  const int s = blockIdx.x, lane = threadIdx.x;
  for (int k = lane; k < nkeys; k += 64) hist[k] = 0;
  __syncthreads();
  const int b = s * seg, e = min(n, b + seg);
  for (int i = b + lane; i < e; i += 64) atomicAdd(&hist[keys[i]], 1);
  __syncthreads();
  for (int k = lane; k < nkeys; k += 64) counts[(size_t)s * nkeys + k] = hist[k];
}

__global__ void __launch_bounds__(256) k_key_totals(const int* __restrict__ counts, int nseg, int nkeys,
                                                    int* __restrict__ totals) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  int t = 0;
  for (int s = 0; s < nseg; s++) t += counts[(size_t)s * nkeys + k];
  totals[k] = t;
}

// exclusive scan of totals[0 .. nkeys) in place, one workgroup of 1024
__global__ void __launch_bounds__(1024) k_key_scan(int* __restrict__ totals, int nkeys) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nkeys + 1023) / 1024, b = t * per, e = min(nkeys, b + per);
  int s = 0;
  for (int k = b; k < e; k++) s += totals[k];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
    const int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t > 0) ? part[t - 1] : 0;
  for (int k = b; k < e; k++) {
    const int v = totals[k];
    totals[k] = run;
    run += v;
  }
}

__global__ void __launch_bounds__(256) k_seg_offsets(const int* __restrict__ counts, const int* __restrict__ start,
                                                     int nseg, int nkeys, int* __restrict__ offsets) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  int run = start[k];
  for (int s = 0; s < nseg; s++) {
    const size_t i = (size_t)s * nkeys + k;
    offsets[i] = run;
    run += counts[i];
  }
}

template <typename V>
__global__ void __launch_bounds__(64) k_seg_scatter(const int* __restrict__ kin, const V* __restrict__ vin, int n,
                                                    int seg, int nkeys, int kbits, const int* __restrict__ offsets,
                                                    int* __restrict__ kout, V* __restrict__ vout) {
  extern __shared__ int base[];
  const int s = blockIdx.x, lane = threadIdx.x;
  for (int k = lane; k < nkeys; k += 64) base[k] = offsets[(size_t)s * nkeys + k];
  __syncthreads();
  const int b = s * seg, e = min(n, b + seg);
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));  // lanes below this one
  for (int i0 = b; i0 < e; i0 += 64) {
    const int i = i0 + lane;
    const bool valid = i < e;
    const int key = valid ? kin[i] : 0;
    // lanes holding the same key: AND of the ballots on each key bit
    unsigned long long same = __ballot(valid);
    for (int bit = 0; bit < kbits; bit++) {
      const unsigned long long on = __ballot(valid && ((key >> bit) & 1));
      same &= ((key >> bit) & 1) ? on : ~on;
    }
    const int rank = __popcll(same & lt);
    if (valid) {
      const int pos = base[key] + rank;
      kout[pos] = key;
      vout[pos] = vin[i];
    }
    __syncthreads();  // every lane has read base[] for this batch
    if (valid && rank == 0) base[key] += __popcll(same);
    __syncthreads();
  }
}

template <typename V>
static int sort_pairs_impl(void* temp, size_t* temp_bytes, const int* kin, int* kout, const V* vin, V* vout, int n,
                           int end_bit, hipStream_t st) {
  if (end_bit < 1 || end_bit > kSortMaxBits || n < 0) return -1;
  const SortPlan p = sort_plan(n, end_bit);
  const size_t cnt = (size_t)p.nseg * p.nkeys;
  const size_t need = (2 * cnt + p.nkeys) * sizeof(int) + 256;
  if (!temp) {
    *temp_bytes = need;
    return 0;
  }
  if (*temp_bytes < need) return -1;
  if (n == 0) return 0;
  int* counts = (int*)temp;
  int* offsets = counts + cnt;
  int* totals = offsets + cnt;
  const size_t lds = (size_t)p.nkeys * sizeof(int);
  hipLaunchKernelGGL(k_seg_count, dim3(p.nseg), dim3(64), lds, st, kin, n, p.seg, p.nkeys, counts);
  hipLaunchKernelGGL(k_key_totals, dim3((p.nkeys + 255) / 256), dim3(256), 0, st, (const int*)counts, p.nseg,
                     p.nkeys, totals);
  hipLaunchKernelGGL(k_key_scan, dim3(1), dim3(1024), 0, st, totals, p.nkeys);
  hipLaunchKernelGGL(k_seg_offsets, dim3((p.nkeys + 255) / 256), dim3(256), 0, st, (const int*)counts,
                     (const int*)totals, p.nseg, p.nkeys, offsets);
  hipLaunchKernelGGL(k_seg_scatter<V>, dim3(p.nseg), dim3(64), lds, st, kin, vin, n, p.seg, p.nkeys, end_bit,
                     (const int*)offsets, kout, vout);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// (key, value) sorts used by the NUFFT set-up; temp == nullptr queries the
// workspace size.  Keys must lie in [0, 2^end_bit).
int sort_pairs(void* temp, size_t* temp_bytes, const int* kin, int* kout, const int* vin, int* vout, int n,
               int end_bit, hipStream_t st) {
  return sort_pairs_impl<int>(temp, temp_bytes, kin, kout, vin, vout, n, end_bit, st);
}
int sort_pairs_u(void* temp, size_t* temp_bytes, const int* kin, int* kout, const unsigned* vin, unsigned* vout,
                 int n, int end_bit, hipStream_t st) {
  return sort_pairs_impl<unsigned>(temp, temp_bytes, kin, kout, vin, vout, n, end_bit, st);
}

}  // namespace mf
