// Dictionary coding of CSR values.
//
// The Hamiltonians of these lattice models have very few distinct matrix
// elements: every entry is a product of a handful of local operator elements
// (0, +-1, +-1/2, +-i/2, Clebsch-Gordan factors ...) times a coupling, and the
// diagonal sums a few local terms.  Storing one uint8/uint16 code per entry
// plus a small table of (value, term) pairs instead of the value (8 or 16 B)
// and the term id (1 B) cuts the streamed bytes of a product from 12-20 to
// 5-6 per entry, and the table is where the per-product term coefficients are
// applied (c[term] * value, once per product instead of once per entry).
//
// Construction, two passes over the entries:
//   1. distinct keys (value bits, term): each block dedups its slice in an
//      LDS hash set and merges it into a global hash set of kDictCap slots.
//      Inside a wave one lane at a time inserts (the lanes holding the same
//      key retire together), so no lane ever waits on a lane of its own wave;
//      a slot being written by another wave or block is waited for.
//   2. the occupied global slots are ranked on the host in slot order, and
//      every entry's code is its key's rank (read-only probing).
// Values are compared bitwise, so a coded product multiplies by exactly the
// stored value.  More than kDictCap distinct keys: the matrix stays uncoded.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace edlgt {
namespace {

constexpr int kDictCap = 8192;  // global hash slots (power of two)
constexpr int kDictMax = 4096;  // largest dictionary kept
constexpr int kLdsSlots = 512;  // per-block hash slots
constexpr int kTB = 256;

struct Key {
  unsigned long long x, y;
  unsigned t;
};

__device__ inline unsigned key_hash(const Key& k) {
  unsigned long long h = k.x * 0x9e3779b97f4a7c15ull ^ (k.y + 0x632be59bd9b4e019ull) * 0xc2b2ae3d27d4eb4full ^
                         (unsigned long long)k.t * 0x165667b19e3779f9ull;
  h ^= h >> 31;
  h *= 0xd6e8feb86659fd93ull;
  h ^= h >> 32;
  return (unsigned)h;
}

template <bool kCplx>
__device__ inline Key entry_key(const double* __restrict__ val, const uint8_t* __restrict__ term, int64_t p) {
  Key k;
  if (kCplx) {
    k.x = (unsigned long long)__double_as_longlong(val[2 * p]);
    k.y = (unsigned long long)__double_as_longlong(val[2 * p + 1]);
  } else {
    k.x = (unsigned long long)__double_as_longlong(val[p]);
    k.y = 0ull;
  }
  k.t = term ? term[p] : 0u;
  return k;
}

// Single-lane insertion into an open-addressing set.  state: 0 empty,
// 1 being written, 2 ready.  Returns false when the set is full.
template <int kScope>
__device__ bool set_insert(const Key& k, unsigned long long* kx, unsigned long long* ky, unsigned* kt,
                           unsigned* state, int cap) {
  const unsigned h = key_hash(k);
  for (int i = 0; i < cap; ++i) {
    const int s = (int)((h + (unsigned)i) & (unsigned)(cap - 1));
    unsigned st = __hip_atomic_load(&state[s], __ATOMIC_ACQUIRE, kScope);
    if (st == 0u) {
      unsigned expect = 0u;
      if (__hip_atomic_compare_exchange_strong(&state[s], &expect, 1u, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                               kScope)) {
        kx[s] = k.x;
        ky[s] = k.y;
        kt[s] = k.t;
        __hip_atomic_store(&state[s], 2u, __ATOMIC_RELEASE, kScope);
        return true;
      }
      st = expect;
    }
    while (st == 1u) st = __hip_atomic_load(&state[s], __ATOMIC_ACQUIRE, kScope);
    if (kx[s] == k.x && ky[s] == k.y && kt[s] == k.t) return true;
  }
  return false;
}

// Wave-cooperative insertion: one lane at a time inserts a key, every lane
// holding that key retires with it.
template <int kScope>
__device__ bool wave_insert(bool pending, const Key& k, unsigned long long* kx, unsigned long long* ky,
                            unsigned* kt, unsigned* state, int cap) {
  bool ok = true;
  const int lane = threadIdx.x & 63;
  while (true) {
    const unsigned long long mask = __ballot(pending);
    if (mask == 0ull) break;
    const int leader = __ffsll((long long)mask) - 1;
    Key l;
    l.x = __shfl(k.x, leader, 64);
    l.y = __shfl(k.y, leader, 64);
    l.t = __shfl(k.t, leader, 64);
    bool ins = true;
    if (lane == leader) ins = set_insert<kScope>(l, kx, ky, kt, state, cap);
    ok = ok && __shfl((int)ins, leader, 64);
    if (pending && k.x == l.x && k.y == l.y && k.t == l.t) pending = false;
  }
  return ok;
}

template <bool kCplx>
__global__ void __launch_bounds__(kTB) dict_collect(const double* __restrict__ val, const uint8_t* __restrict__ term,
                                                    int64_t nnz, unsigned long long* gx, unsigned long long* gy,
                                                    unsigned* gt, unsigned* gstate, int* overflow) {
  __shared__ unsigned long long lx[kLdsSlots], ly[kLdsSlots];
  __shared__ unsigned lt[kLdsSlots], lstate[kLdsSlots];
  for (int i = threadIdx.x; i < kLdsSlots; i += kTB) lstate[i] = 0u;
  __syncthreads();
  const int64_t chunk = (nnz + gridDim.x - 1) / gridDim.x;
  const int64_t p0 = (int64_t)blockIdx.x * chunk, p1 = min(nnz, p0 + chunk);
  bool gfail = false;
  for (int64_t base = p0; base < p1; base += kTB) {
    const int64_t p = base + threadIdx.x;
    const bool have = p < p1;
    Key k{0ull, 0ull, 0u};
    if (have) k = entry_key<kCplx>(val, term, p);
    // fast path: already in the block's set (read-only probe of ready slots)
    bool pending = have;
    if (pending) {
      const unsigned h = key_hash(k);
      for (int i = 0; i < kLdsSlots; ++i) {
        const int s = (int)((h + (unsigned)i) & (kLdsSlots - 1));
        const unsigned st = __hip_atomic_load(&lstate[s], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (st != 2u) break;
        if (lx[s] == k.x && ly[s] == k.y && lt[s] == k.t) {
          pending = false;
          break;
        }
      }
    }
    const bool ok = wave_insert<__HIP_MEMORY_SCOPE_WORKGROUP>(pending, k, lx, ly, lt, lstate, kLdsSlots);
    if (!ok) {
      // block set full: this batch goes straight to the global set
      if (!wave_insert<__HIP_MEMORY_SCOPE_AGENT>(have, k, gx, gy, gt, gstate, kDictCap)) gfail = true;
    }
  }
  __syncthreads();
  // merge the block's set into the global set
  for (int s0 = 0; s0 < kLdsSlots; s0 += kTB) {
    const int s = s0 + threadIdx.x;
    const bool have = lstate[s] == 2u;
    Key k{0ull, 0ull, 0u};
    if (have) k = Key{lx[s], ly[s], lt[s]};
    if (!wave_insert<__HIP_MEMORY_SCOPE_AGENT>(have, k, gx, gy, gt, gstate, kDictCap)) gfail = true;
  }
  if (gfail) atomicOr(overflow, 1);
}

template <bool kCplx, class CT>
__global__ void __launch_bounds__(kTB) dict_encode(const double* __restrict__ val, const uint8_t* __restrict__ term,
                                                   int64_t nnz, const unsigned long long* __restrict__ gx,
                                                   const unsigned long long* __restrict__ gy,
                                                   const unsigned* __restrict__ gt,
                                                   const int* __restrict__ rank, CT* __restrict__ code,
                                                   int* __restrict__ missing) {
  for (int64_t p = (int64_t)blockIdx.x * kTB + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * kTB) {
    const Key k = entry_key<kCplx>(val, term, p);
    const unsigned h = key_hash(k);
    int found = -1;
    for (int i = 0; i < kDictCap; ++i) {
      const int s = (int)((h + (unsigned)i) & (kDictCap - 1));
      if (rank[s] < 0) break;
      if (gx[s] == k.x && gy[s] == k.y && gt[s] == k.t) {
        found = rank[s];
        break;
      }
    }
    if (found < 0) {
      atomicOr(missing, 1);
      found = 0;
    }
    code[p] = (CT)found;
  }
}

}  // namespace

int csr_compress(Csr& h, hipStream_t st, int* compressed) {
  *compressed = 0;
  if (h.n_dict > 0) {
    *compressed = 1;
    return EDLGT_OK;
  }
  if (h.value_type != EDLGT_VALUE_F64 && h.value_type != EDLGT_VALUE_C128)
    return set_error(EDLGT_ERR_INVALID_ARG, "only FP64 matrices are coded");
  const bool cplx = h.value_type == EDLGT_VALUE_C128;
  const int64_t nnz = h.nnz;
  DevBuf gx, gy, gt, gstate, flags, rank;
  EDLGT_TRY(gx.alloc(kDictCap * sizeof(unsigned long long)));
  EDLGT_TRY(gy.alloc(kDictCap * sizeof(unsigned long long)));
  EDLGT_TRY(gt.alloc(kDictCap * sizeof(unsigned)));
  EDLGT_TRY(gstate.alloc(kDictCap * sizeof(unsigned)));
  EDLGT_TRY(flags.alloc(2 * sizeof(int)));
  EDLGT_TRY(rank.alloc(kDictCap * sizeof(int)));
  EDLGT_HIP(hipMemsetAsync(gstate.p, 0, kDictCap * sizeof(unsigned), st));
  EDLGT_HIP(hipMemsetAsync(flags.p, 0, 2 * sizeof(int), st));
  const uint8_t* term = h.n_terms > 1 ? h.term.as<uint8_t>() : nullptr;
  const double* val = h.val.as<double>();
  if (nnz > 0) {
    const unsigned nblk = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (nnz + 65535) / 65536));
    if (cplx)
      hipLaunchKernelGGL(dict_collect<true>, dim3(nblk), dim3(kTB), 0, st, val, term, nnz,
                         gx.as<unsigned long long>(), gy.as<unsigned long long>(), gt.as<unsigned>(),
                         gstate.as<unsigned>(), flags.as<int>());
    else
      hipLaunchKernelGGL(dict_collect<false>, dim3(nblk), dim3(kTB), 0, st, val, term, nnz,
                         gx.as<unsigned long long>(), gy.as<unsigned long long>(), gt.as<unsigned>(),
                         gstate.as<unsigned>(), flags.as<int>());
    EDLGT_HIP(hipGetLastError());
  }
  std::vector<unsigned> hstate(kDictCap), ht(kDictCap);
  std::vector<unsigned long long> hx(kDictCap), hy(kDictCap);
  int hflags[2] = {0, 0};
  EDLGT_HIP(hipMemcpyAsync(hstate.data(), gstate.p, kDictCap * sizeof(unsigned), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipMemcpyAsync(hx.data(), gx.p, kDictCap * 8, hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipMemcpyAsync(hy.data(), gy.p, kDictCap * 8, hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipMemcpyAsync(ht.data(), gt.p, kDictCap * sizeof(unsigned), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipMemcpyAsync(hflags, flags.p, sizeof(int), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  if (hflags[0]) return EDLGT_OK;  // more than kDictCap distinct keys: stays uncoded
  std::vector<int> hrank(kDictCap, -1);
  std::vector<double2> dval;
  std::vector<int32_t> dterm;
  for (int s = 0; s < kDictCap; ++s) {
    if (hstate[s] != 2u) continue;
    hrank[s] = (int)dval.size();
    double a, b;
    std::memcpy(&a, &hx[s], 8);
    std::memcpy(&b, &hy[s], 8);
    dval.push_back(make_double2(a, cplx ? b : 0.0));
    dterm.push_back((int32_t)ht[s]);
  }
  const int nd = (int)dval.size();
  if (nd > kDictMax) return EDLGT_OK;
  const int cb = nd <= 256 ? 1 : 2;
  DevBuf code, dv, dt;
  EDLGT_TRY(code.alloc((size_t)std::max<int64_t>(nnz, 1) * cb));
  EDLGT_TRY(dv.alloc((size_t)std::max(nd, 1) * sizeof(double2)));
  EDLGT_TRY(dt.alloc((size_t)std::max(nd, 1) * sizeof(int32_t)));
  if (nd > 0) {
    EDLGT_HIP(hipMemcpyAsync(dv.p, dval.data(), nd * sizeof(double2), hipMemcpyHostToDevice, st));
    EDLGT_HIP(hipMemcpyAsync(dt.p, dterm.data(), nd * sizeof(int32_t), hipMemcpyHostToDevice, st));
  }
  EDLGT_HIP(hipMemcpyAsync(rank.p, hrank.data(), kDictCap * sizeof(int), hipMemcpyHostToDevice, st));
  if (nnz > 0) {
    const unsigned g = grid_for(nnz, kTB, 65536);
#define ENC(CPLX, CT)                                                                                   \
  hipLaunchKernelGGL((dict_encode<CPLX, CT>), dim3(g), dim3(kTB), 0, st, val, term, nnz,                \
                     gx.as<unsigned long long>(), gy.as<unsigned long long>(), gt.as<unsigned>(),      \
                     rank.as<int>(), code.as<CT>(), flags.as<int>() + 1)
    if (cplx && cb == 1) ENC(true, uint8_t);
    else if (cplx) ENC(true, uint16_t);
    else if (cb == 1) ENC(false, uint8_t);
    else ENC(false, uint16_t);
#undef ENC
    EDLGT_HIP(hipGetLastError());
  }
  EDLGT_HIP(hipMemcpyAsync(hflags + 1, flags.as<int>() + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  EDLGT_HIP(hipStreamSynchronize(st));
  if (hflags[1]) return set_error(EDLGT_ERR_HIP, "dictionary lookup missed a key");
  DevBuf tu;
  if (cplx) {
    EDLGT_TRY(upload(tu, dval.data(), std::max<size_t>(dval.size(), 1), st));
  } else {
    std::vector<double> re(std::max<size_t>(dval.size(), 1), 0.0);
    for (size_t i = 0; i < dval.size(); ++i) re[i] = dval[i].x;
    EDLGT_TRY(upload(tu, re.data(), re.size(), st));
  }
  EDLGT_HIP(hipStreamSynchronize(st));
  h.tab_unit = std::move(tu);
  h.code = std::move(code);
  h.dict_val = std::move(dv);
  h.dict_term = std::move(dt);
  h.n_dict = nd;
  h.code_bytes = cb;
  *compressed = 1;
  return EDLGT_OK;
}

}  // namespace edlgt
