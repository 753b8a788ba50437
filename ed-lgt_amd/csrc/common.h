// Shared internals of libedlgt_hip: status codes, error capture, device
// buffers.  Everything here is internal; the public surface is include/edlgt_hip.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <memory>
#include <string>
#include <vector>

#include "../../include/edlgt_hip.h"

namespace edlgt {

typedef unsigned __int128 u128;

// Records a message for edlgt_last_error() and returns the code.
int set_error(int code, const std::string& msg);
inline int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return EDLGT_OK;
  return set_error(e == hipErrorOutOfMemory ? EDLGT_ERR_OUT_OF_MEMORY : EDLGT_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

#define EDLGT_HIP(call)                                     \
  do {                                                      \
    hipError_t _e = (call);                                 \
    if (_e != hipSuccess) return ::edlgt::hip_status(_e, #call); \
  } while (0)

#define EDLGT_TRY(call)                 \
  do {                                  \
    int _s = (call);                    \
    if (_s != EDLGT_OK) return _s;      \
  } while (0)

// Owning device allocation (hipMalloc/hipFree), movable, not copyable.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return set_error(EDLGT_ERR_OUT_OF_MEMORY,
                       "hipMalloc of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e));
    }
    bytes = n;
    return EDLGT_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
int upload(DevBuf& b, const T* host, size_t count, hipStream_t s) {
  EDLGT_TRY(b.alloc(count * sizeof(T)));
  if (count) EDLGT_HIP(hipMemcpyAsync(b.p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
  return EDLGT_OK;
}

// Device-wide exclusive scan of int32 counts into int64 offsets, with the total
// written to offsets[n].  Implemented in scan.hip.
int exclusive_scan_i32_i64(const int32_t* counts, int64_t* offsets, int64_t n, hipStream_t s);
int exclusive_scan_i64(const int64_t* counts, int64_t* offsets, int64_t n, hipStream_t s);

// Sorted sector basis kept on the device.
struct Basis {
  int32_t n_sites = 0;
  int32_t cfg_stride = 0;     // bytes per configuration row (n_sites rounded up to 4)
  int64_t n_states = 0;
  std::vector<int32_t> loc_dims;
  std::vector<uint64_t> strides;  // mixed-radix, last site fastest (64-bit keys)
  std::vector<u128> strides128;   // same strides for 128-bit keys
  int32_t key_words = 1;      // 1: uint64 keys; 2: unsigned 128-bit keys (prod(loc_dims) >= 2^64)
  DevBuf keys;                // key_words * 8 bytes per state, ascending
  DevBuf configs;             // uint8  [n_states][cfg_stride]
};

// On-the-fly Hamiltonian (term program over the sector basis), defined in
// assemble.hip.
struct MatFree;

// CSR Hamiltonian kept on the device.
struct Csr {
  int64_t n_rows = 0;
  int64_t nnz = 0;
  int32_t value_type = EDLGT_VALUE_C128;
  int32_t g_override = 0;     // lanes per row of the SpMV kernels (0: from the mean row length)
  int64_t grid_override = 0;  // SpMV workgroups (0: default), set by the autotuner
  DevBuf row_ptr;   // int64 [n_rows + 1]
  DevBuf col;       // int32 [nnz]
  DevBuf val;       // value_type [nnz]
  DevBuf val_lo;    // reduced-precision copy (complex64 / float32), built on demand
  DevBuf term;      // uint8 [nnz] term id of every entry (multi-term matrices only)
  int32_t n_terms = 1;
  // CSR-stream tiles of 1024 / 2048 / 4096 nonzeros (tile_blocks[t]: first
  // row of every tile, tile_pos[t]: its first entry); 2048 is the default
  DevBuf tile_blocks[3], tile_pos[3];
  int64_t n_tile_blocks[3] = {0, 0, 0};
  DevBuf& stream_blocks = tile_blocks[1];
  DevBuf& stream_pos = tile_pos[1];
  int64_t n_stream_blocks = 0;
  int32_t coded_variant = 0;  // launch variant of the coded stream kernel
  int32_t plain_variant = 0;  // launch variant of the uncoded stream kernel
  // dictionary coding of the values (edlgt_csr_compress): entry p holds
  // dict_val[code[p]] of term dict_term[code[p]]; code is uint8 when
  // n_dict <= 256, else uint16 (n_dict <= 4096)
  int32_t n_dict = 0;
  int32_t code_bytes = 0;
  DevBuf code;
  DevBuf dict_val;   // double2 [n_dict] (imaginary part 0 for real matrices)
  DevBuf dict_term;  // int32 [n_dict]
  DevBuf tab_unit;   // dict values with all term coefficients 1 (double or double2)
  // matrix-free operator: no stored entries, every product generates them
  // from the term program (nnz = 0, row_ptr / col / val empty)
  std::shared_ptr<MatFree> mf;
};

// Momentum-sector projector B (n_rows x n_cols): CSC (col_ptr, row_idx,
// data) and, per sector row, its column (-1: in no kept column) and amplitude.
struct Momentum {
  int64_t n_rows = 0, n_cols = 0, nnz = 0;
  int32_t is_complex = 0;  // 1 for k != 0 (complex Bloch phases)
  DevBuf col_ptr;  // int64 [n_cols + 1]
  DevBuf row_idx;  // int32 [nnz], ascending within a column
  DevBuf data;     // double2 [nnz]
  DevBuf row_col;  // int32 [n_rows]
  DevBuf row_val;  // double2 [n_rows]
};

inline size_t value_bytes(int32_t vt) {
  switch (vt) {
    case EDLGT_VALUE_F64: return 8;
    case EDLGT_VALUE_C128: return 16;
    case EDLGT_VALUE_F32: return 4;
    case EDLGT_VALUE_C64: return 8;
  }
  return 0;
}

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace edlgt

struct edlgt_basis : edlgt::Basis {};
struct edlgt_csr : edlgt::Csr {};
struct edlgt_momentum : edlgt::Momentum {};
