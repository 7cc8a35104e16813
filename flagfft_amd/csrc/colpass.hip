// colpass.hip — register-resident two-stage super-stages (gfx950): the
// standalone kernels and their catalogue.  The pass body (colpass_body.hpp)
// is shared with the persistent queue kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "colpass_body.hpp"
#include "kernels.hpp"

namespace flagfft {
namespace dev {

// WPE: waves per SIMD the register allocation must leave room for (two
// workgroups of C * RB threads per CU: C * RB / 128).
// PAIR: the instance with the DC / Nyquist pair modes of a 2D real column
// pass (a separate instance keeps that code out of every other pass).
template <typename T, int C, int RA, int RB, int NCH, int WPE, int TWM, bool PAIR>
__global__ __launch_bounds__(C * RB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_col(const StageArgs a) {
  col_body<T, C, RA, RB, NCH, TWM, PAIR>(a, remap_block(blockIdx.x, a.nblocks, a.xcd_remap));
}

}  // namespace dev

namespace {

struct ColEntry {
  int r, ra, rb, c, nch, wpe;
  bool dbl;
  const void *f;
  int pref;  // default instance for first passes (1), later passes (2), both (3), neither (0)
  int qvar;  // the persistent queue kernel's variant of this instance (k_queue in kernels.hip), -1 if none
  const void *fp;  // the instance with the DC / Nyquist pair modes, nullptr if none
};

#define FLAGFFT_COL(T, C, RA, RB, NCH, WPE, PREF)                                                                  \
  {RA * RB, RA, RB, C, NCH, WPE, sizeof(T) == 8, (const void *)dev::k_col<T, C, RA, RB, NCH, WPE, 1, false>, PREF, \
   -1, nullptr}
#define FLAGFFT_COLP(T, C, RA, RB, NCH, WPE, PREF)                                                                 \
  {RA * RB, RA, RB, C, NCH, WPE, sizeof(T) == 8, (const void *)dev::k_col<T, C, RA, RB, NCH, WPE, 1, false>, PREF, \
   -1, (const void *)dev::k_col<T, C, RA, RB, NCH, WPE, 1, true>}
#define FLAGFFT_COLQ(T, C, RA, RB, NCH, WPE, PREF, QVAR)                                                           \
  {RA * RB, RA, RB, C, NCH, WPE, sizeof(T) == 8, (const void *)dev::k_col<T, C, RA, RB, NCH, WPE, 1, false>, PREF, \
   QVAR, nullptr}

// Measured on MI355X (2^20 x 16 and 2^24 C2C FP32): for R = 4096 (64 x 64
// in registers, 198 VGPRs, one 512-thread workgroup per CU) 8 units per
// workgroup, which makes 2^24 two passes instead of three (150 against
// 185 us) and the 4096-point columns of a 4096^2 transform one pass instead
// of two (75 against 95 us); for R = 1024 the first
// pass (contiguous stores) runs best with 32 units per workgroup, later
// passes (C-element stores) with 16; for R = 256 the first pass with 16
// units (65-67 us per 2^24 pass against 72-73 with 64), later passes with 64
// units in two exchange chunks (55-59 us against 62-65 with 16).
const ColEntry kColTable[] = {
    FLAGFFT_COLP(float, 8, 64, 64, 2, 2, 3),          // 4096: 64-byte runs, 135 KB LDS, 198 VGPRs
    FLAGFFT_COL(float, 16, 32, 32, 2, 4, 2),          // 1024: 128-byte runs, 68 KB LDS, two workgroups per CU
    FLAGFFT_COL(float, 32, 32, 32, 2, 4, 1),          // 1024: 256-byte runs, 136 KB LDS, 1024 threads
    FLAGFFT_COL(float, 16, 32, 32, 1, 2, 0),          // 1024: 132 KB LDS, one workgroup per CU
    FLAGFFT_COLQ(float, 8, 32, 32, 2, 4, 0, 0),       // 1024: 64-byte runs, 35 KB LDS, four workgroups per CU
    FLAGFFT_COL(float, 8, 32, 32, 1, 2, 0),           // 1024: 64-byte runs
    FLAGFFT_COLQ(float, 16, 16, 16, 1, 2, 1, 2),      // 256: 128-byte runs, 33 KB LDS
    FLAGFFT_COL(float, 32, 16, 16, 1, 4, 0),          // 256: 256-byte runs, 66 KB LDS
    FLAGFFT_COL(float, 64, 16, 16, 2, 4, 2),          // 256: 512-byte runs, 66 KB LDS, 1024 threads
    FLAGFFT_COLQ(float, 16, 32, 16, 1, 2, 3, 1),      // 512: 128-byte runs
    FLAGFFT_COLQ(float, 32, 16, 8, 1, 2, 3, 3),       // 128
    FLAGFFT_COL(float, 64, 8, 8, 1, 4, 3),            // 64: 512-byte runs
    FLAGFFT_COLQ(float, 32, 8, 8, 1, 2, 0, 4),        // 64: 256-byte runs
    FLAGFFT_COLQ(float, 64, 8, 4, 1, 2, 3, 5),        // 32: 512-byte runs
    FLAGFFT_COLQ(float, 64, 4, 4, 1, 2, 3, 6),        // 16: 512-byte runs
    FLAGFFT_COLQ(double, 16, 16, 16, 1, 2, 3, 7),     // 256: 256-byte runs, 66 KB LDS
    FLAGFFT_COL(double, 8, 16, 16, 1, 2, 0),          // 256: 128-byte runs
    FLAGFFT_COL(double, 16, 16, 8, 1, 2, 3),          // 128
    FLAGFFT_COLQ(double, 32, 8, 8, 1, 2, 3, 9),       // 64
};

}  // namespace

// units: 0 for the default instance of a first (first = true) or later
// pass, C for the first instance with C units per workgroup, 1000 NCH + C
// for the one with NCH exchange chunks as well.
int col_kernel_id(int64_t r, bool dbl, int units, bool first) {
  const int n = static_cast<int>(sizeof(kColTable) / sizeof(kColTable[0]));
  const int c = units % 1000, nch = units / 1000;
  const int want = first ? 1 : 2;
  for (int id = 0; id < n; ++id) {
    const ColEntry &e = kColTable[id];
    if (e.r != r || e.dbl != dbl) continue;
    if (units <= 0 ? (e.pref & want) != 0 : (e.c == c && (nch == 0 || e.nch == nch))) return id;
  }
  return -1;
}

int col_kernel_id(int64_t r, bool dbl, int units) { return col_kernel_id(r, dbl, units, false); }

int col_run_bytes(int64_t r, bool dbl, bool first) {
  const int id = col_kernel_id(r, dbl, 0, first);
  return id < 0 ? 0 : kColTable[id].c * (dbl ? 16 : 8);
}

int col_kernel_id_dividing(int64_t r, bool dbl, int64_t tcnt) {
  const int n = static_cast<int>(sizeof(kColTable) / sizeof(kColTable[0]));
  int any = -1;
  for (int id = 0; id < n; ++id) {
    const ColEntry &e = kColTable[id];
    if (e.r != r || e.dbl != dbl || tcnt % e.c != 0) continue;
    if (e.pref != 0) return id;
    if (any < 0) any = id;
  }
  return any;
}

int col_queue_kernel_id(int64_t r, bool dbl) {
  const int n = static_cast<int>(sizeof(kColTable) / sizeof(kColTable[0]));
  for (int id = 0; id < n; ++id)
    if (kColTable[id].r == r && kColTable[id].dbl == dbl && kColTable[id].qvar >= 0) return id;
  return -1;
}

int col_queue_variant(int id) { return kColTable[id].qvar; }

void col_kernel_shape(int id, int &units, int &threads, size_t &lds_bytes, int &elem_bytes) {
  const ColEntry &e = kColTable[id];
  units = e.c;
  threads = e.c * e.rb;
  elem_bytes = e.dbl ? 16 : 8;
  int need = 0;
  if (e.nch == 2 && e.ra == e.rb) {
    need = e.ra * ((e.ra / 2) * e.c + 8);  // two-chunk square form: RA rows of (RA / 2) C + 8 entries
  } else {
    const int ex = (e.ra / e.nch) * e.rb * e.c;                 // exchange chunk entries (+ 1 per 32)
    const int bl = e.r / e.nch;
    const int blk = e.c * ((bl + bl / 32 + 28) / 32 * 32 + 4);  // block-store chunk entries
    need = std::max(ex + (ex >> 5) + 1, blk);
  }
  lds_bytes = static_cast<size_t>(need) * static_cast<size_t>(elem_bytes);
}

const void *col_kernel(int id) { return kColTable[id].f; }

const void *col_pair_kernel(int id) { return kColTable[id].fp; }

int col_pair_kernel_id(int64_t r, bool dbl) {
  const int id = col_kernel_id(r, dbl, 0, true);
  return id >= 0 && kColTable[id].fp != nullptr ? id : -1;
}

std::string col_kernel_targs(int id, bool pair) {
  const ColEntry &e = kColTable[id];
  return std::to_string(e.c) + "," + std::to_string(e.ra) + "," + std::to_string(e.rb) + "," + std::to_string(e.nch) +
         "," + std::to_string(e.wpe) + ",1," + (pair ? "true" : "false");
}

}  // namespace flagfft
