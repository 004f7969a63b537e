// Device helpers shared by the streaming steppers (one wavefront per z
// column, CPL consecutive z cells per lane).
#pragma once

#include "fdtd_common.h"

namespace pjz {

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Global (address space 1) views of device pointers.  Accesses through them
// compile to global loads and stores, not flat ones.  A kernel whose
// pointers come from an argument block in memory (the batched kernels) would
// otherwise access memory with flat instructions, which count in lgkmcnt as
// well as vmcnt: every LDS barrier (lgkmcnt(0)) would then wait for the
// prefetched loads and the stores in flight.  A pending flat operation also
// makes the compiler wait for every outstanding vector memory operation
// (vmcnt(0)) at the next join.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const float gcfloat;
typedef __attribute__((address_space(1))) f32x2 gf32x2;
typedef __attribute__((address_space(1))) const f32x2 gcf32x2;
typedef __attribute__((address_space(1))) f32x4 gf32x4;
typedef __attribute__((address_space(1))) const f32x4 gcf32x4;
__device__ inline gfloat* to_global(float* p) { return (gfloat*)p; }
__device__ inline gcfloat* to_global(const float* p) { return (gcfloat*)p; }

// Loads CPL consecutive z cells of a row; cells at z >= Z read as zero.  NT
// marks the load nontemporal (streaming: the line is not kept for reuse).
template <int CPL, bool vec, bool NT = false>
__device__ inline void load_col(float (&v)[CPL], const float* __restrict__ pg, int zb, int Z,
                                bool valid) {
  if (!valid || zb >= Z) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) v[c] = 0.0f;
    return;
  }
  gcfloat* p = to_global(pg);
  if (vec) {
    if constexpr (CPL == 4) {
      gcf32x4* q = reinterpret_cast<gcf32x4*>(p + zb);
      const f32x4 t = NT ? __builtin_nontemporal_load(q) : *q;
      v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else if constexpr (CPL == 2) {
      gcf32x2* q = reinterpret_cast<gcf32x2*>(p + zb);
      const f32x2 t = NT ? __builtin_nontemporal_load(q) : *q;
      v[0] = t.x; v[1] = t.y;
    } else if constexpr (CPL == 3) {
      // three 4-byte-aligned floats: merged into one 96-bit load
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = NT ? __builtin_nontemporal_load(p + zb + c) : p[zb + c];
    } else {
      v[0] = NT ? __builtin_nontemporal_load(p + zb) : p[zb];
    }
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) v[c] = (zb + c < Z) ? p[zb + c] : 0.0f;
  }
}

// Stores CPL consecutive z cells (z < Z only).  Field results are written
// nontemporally: no kernel re-reads them before the next launch, and keeping
// them out of the cache leaves room for the inputs that neighbouring tiles
// share (ghost rows).
template <int CPL, bool vec>
__device__ inline void store_col(float* __restrict__ pg, const float (&v)[CPL], int zb, int Z) {
  if (zb >= Z) return;
  gfloat* p = to_global(pg);
  if (vec) {
    if constexpr (CPL == 4) {
      f32x4 t = {v[0], v[1], v[2], v[3]};
      __builtin_nontemporal_store(t, reinterpret_cast<gf32x4*>(p + zb));
    } else if constexpr (CPL == 2) {
      f32x2 t = {v[0], v[1]};
      __builtin_nontemporal_store(t, reinterpret_cast<gf32x2*>(p + zb));
    } else if constexpr (CPL == 3) {
#pragma unroll
      for (int c = 0; c < 3; ++c) __builtin_nontemporal_store(v[c], p + zb + c);
    } else {
      __builtin_nontemporal_store(v[0], p + zb);
    }
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c)
      if (zb + c < Z) p[zb + c] = v[c];
  }
}

// store_col through a global pointer.
template <int CPL, bool vec>
__device__ inline void store_col_global(gfloat* __restrict__ p, const float (&v)[CPL], int zb,
                                        int Z) {
  if (zb >= Z) return;
  if (vec) {
    if constexpr (CPL == 4) {
      f32x4 t = {v[0], v[1], v[2], v[3]};
      __builtin_nontemporal_store(t, (gf32x4*)(p + zb));
    } else if constexpr (CPL == 2) {
      f32x2 t = {v[0], v[1]};
      __builtin_nontemporal_store(t, (gf32x2*)(p + zb));
    } else if constexpr (CPL == 3) {
#pragma unroll
      for (int c = 0; c < 3; ++c) __builtin_nontemporal_store(v[c], p + zb + c);
    } else {
      __builtin_nontemporal_store(v[0], p + zb);
    }
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c)
      if (zb + c < Z) p[zb + c] = v[c];
  }
}

// Raw buffer access: a 128-bit resource (base, size in bytes) in scalar
// registers plus 32-bit byte offsets, one vector (VGPR) and one uniform
// (SGPR).  An offset at or past the resource's size reads 0 and drops the
// store, so kOff replaces per-lane bounds branches.  Resources here are at
// most 2^31 bytes, which kOff is never below.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr unsigned kOff = 0x80000000u;
// Resource word 3 of a raw (untyped, unswizzled) buffer on gfx9-family GPUs.
constexpr int kRsrcWord3 = 0x00020000;
// Cache-policy operand of buffer stores: nontemporal (the nt bit on gfx950),
// for the same reason as store_col's.
constexpr int kStoreNT = 2;

__device__ inline rsrc_t make_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           kRsrcWord3);
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline float bload1(rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

__device__ inline void bstore1(float v, rsrc_t r, unsigned voff, unsigned soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, kStoreNT);
}

// CPL consecutive floats at byte offset voff + soff.
template <int CPL>
__device__ inline void bload(float (&v)[CPL], rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (CPL == 4) {
    const u32x4 t = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    v[0] = __builtin_bit_cast(float, t.x);
    v[1] = __builtin_bit_cast(float, t.y);
    v[2] = __builtin_bit_cast(float, t.z);
    v[3] = __builtin_bit_cast(float, t.w);
  } else if constexpr (CPL == 3) {
    const u32x3 t = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
    v[0] = __builtin_bit_cast(float, t.x);
    v[1] = __builtin_bit_cast(float, t.y);
    v[2] = __builtin_bit_cast(float, t.z);
  } else if constexpr (CPL == 2) {
    const u32x2 t = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    v[0] = __builtin_bit_cast(float, t.x);
    v[1] = __builtin_bit_cast(float, t.y);
  } else {
    v[0] = bload1(r, voff, soff);
  }
}

template <int CPL>
__device__ inline void bstore(const float (&v)[CPL], rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (CPL == 4) {
    u32x4 t;
    t.x = __builtin_bit_cast(unsigned, v[0]);
    t.y = __builtin_bit_cast(unsigned, v[1]);
    t.z = __builtin_bit_cast(unsigned, v[2]);
    t.w = __builtin_bit_cast(unsigned, v[3]);
    __builtin_amdgcn_raw_buffer_store_b128(t, r, voff, soff, kStoreNT);
  } else if constexpr (CPL == 3) {
    u32x3 t;
    t.x = __builtin_bit_cast(unsigned, v[0]);
    t.y = __builtin_bit_cast(unsigned, v[1]);
    t.z = __builtin_bit_cast(unsigned, v[2]);
    __builtin_amdgcn_raw_buffer_store_b96(t, r, voff, soff, kStoreNT);
  } else if constexpr (CPL == 2) {
    u32x2 t;
    t.x = __builtin_bit_cast(unsigned, v[0]);
    t.y = __builtin_bit_cast(unsigned, v[1]);
    __builtin_amdgcn_raw_buffer_store_b64(t, r, voff, soff, kStoreNT);
  } else {
    bstore1(v[0], r, voff, soff);
  }
}

// Read-only data at a wave-uniform index, loaded through the scalar cache
// (the constant address space lets the compiler use a scalar load), which
// keeps it off the vector memory pipeline.
typedef const __attribute__((address_space(4))) float* const_fptr;
__device__ inline float uniform_load(const float* p, int i) { return ((const_fptr)p)[i]; }

// Lane l+1's value (lane_next) or lane l-1's (lane_prev) across the whole
// wave, 0 past its ends: DPP moves (wave_shl:1, wave_shr:1), plain VALU
// instructions.  A shuffle would go through the LDS crossbar (ds_bpermute),
// which competes with the kernels' LDS traffic and has to be waited for.
__device__ inline float lane_next(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
}
__device__ inline float lane_prev(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}

// Cell z+1 for each of the lane's cells (0 at and beyond Z).  A z column
// spans W lanes (64, or 32 when a wave holds two rows).
template <int CPL, int W = 64>
__device__ inline void z_next(float (&o)[CPL], const float (&v)[CPL], int zb, int Z) {
  float up;
  if constexpr (W == 64) up = lane_next(v[0]);
  else up = __shfl_down(v[0], 1, W);
#pragma unroll
  for (int c = 0; c < CPL - 1; ++c) o[c] = v[c + 1];
  o[CPL - 1] = up;
#pragma unroll
  for (int c = 0; c < CPL; ++c)
    if (zb + c + 1 >= Z) o[c] = 0.0f;
}

// Cell z-1 for each of the lane's cells (0 below z = 0); `lane` counts
// within the column's W lanes.
template <int CPL, int W = 64>
__device__ inline void z_prev(float (&o)[CPL], const float (&v)[CPL], int lane) {
  float dn;
  if constexpr (W == 64) dn = lane_prev(v[CPL - 1]);
  else dn = __shfl_up(v[CPL - 1], 1, W);
  o[0] = lane == 0 ? 0.0f : dn;
#pragma unroll
  for (int c = 1; c < CPL; ++c) o[c] = v[c - 1];
}

// Copies from global memory straight into LDS (gfx950 global_load_lds):
// the LDS address goes in M0 and is wave-uniform; lane i writes its bytes
// at lds + i * size.  Issued by inline asm, so the compiler neither counts
// nor waits for them: a kernel using them waits itself (s_waitcnt vmcnt)
// before reading what they wrote.  The compiler does not preserve M0 across
// an asm statement that clobbers it, so a kernel using them must make no
// other use of M0 (no indirect register indexing, messages or LDS copies of
// the compiler's own); tests/test_lds_dma_isa.py checks the built code.
typedef const __attribute__((address_space(3))) void* lds_cptr;

__device__ inline unsigned lds_addr(const void* p) { return (unsigned)(size_t)(lds_cptr)p; }

// 16 bytes a lane.
__device__ inline void dma16(const float* g, unsigned lds) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(lds)
               : "memory", "m0");
}

// 16 bytes a lane, nontemporal (data read once).
__device__ inline void dma16nt(const float* g, unsigned lds) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off nt" ::"v"(g), "s"(lds)
               : "memory", "m0");
}

// 4 bytes a lane.
__device__ inline void dma4(const float* g, unsigned lds) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dword %0, off" ::"v"(g), "s"(lds)
               : "memory", "m0");
}

// Workgroup barrier that waits for this wave's LDS (and scalar) operations
// only.  __syncthreads() also waits for every outstanding global load and
// store (vmcnt(0)), which would drain the prefetched plane at every barrier.
// Rows exchange data only through LDS, so LDS ordering is all that is needed.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Adds J(n) of the source plane to the curl terms of cell (x, y, z) (the
// update is E = ca E + cb (curl H + J)).  The two source components land in
// fixed slots (ju, jv) on every branch and are added with selects on the
// (uniform) axis.  Adding into cx, cy or cz on different branches lets the
// compiler merge the branches into a store through a selected pointer: the
// curl terms then live in scratch memory, and the wait for that scratch load
// also waits for every prefetched load of the next plane.
__device__ inline void add_source(const Src& src, const Geom& g, int n, int x, int y, int z,
                                  float& cx, float& cy, float& cz) {
  float ju = 0.0f, jv = 0.0f;  // J of (y, z), (x, z) or (x, y) for axis 0, 1, 2
  if (z < g.Z) {
    if (src.axis == 0) {
      if (x == src.pos || x == src.pos - 1) {
        const float wv = src.wave[2 * n + (x == src.pos ? 0 : 1)];
        const long long so = (long long)y * g.Z + z;
        ju = src.field[so] * wv;
        jv = src.field[(long long)g.Y * g.Z + so] * wv;
      }
    } else if (src.axis == 1) {
      if (y == src.pos || y == src.pos - 1) {
        const float wv = src.wave[2 * n + (y == src.pos ? 0 : 1)];
        const long long so = (long long)x * g.Z + z;
        ju = src.field[so] * wv;
        jv = src.field[(long long)g.X * g.Z + so] * wv;
      }
    } else if (z == src.pos) {
      const long long xy = (long long)x * g.Y + y, XY = (long long)g.X * g.Y;
      for (int q = 0; q < src.quad; ++q) {
        const float wv = src.wave[2 * n + q];
        ju += src.field[(2 * q + 0) * XY + xy] * wv;
        jv += src.field[(2 * q + 1) * XY + xy] * wv;
      }
    }
  }
  cx += src.axis == 0 ? 0.0f : ju;
  cy += src.axis == 0 ? ju : (src.axis == 1 ? 0.0f : jv);
  cz += src.axis == 2 ? 0.0f : jv;
}

// Work queue of the persistent batch kernels.  A launch's items [0, total)
// (simulation by simulation, a simulation's row tiles of one x chunk
// consecutive) are split into kQueues contiguous ranges, one per XCD.  A
// workgroup takes items from the range of the XCD it runs on (the XCC_ID
// hardware register) first, so y-adjacent tiles, which share ghost rows, run
// under one L2 cache, and then from the other ranges once its own is empty.
// (Measured against one shared counter: batch64 1000 steps 6.18 -> 5.83 s,
// design_batch forward + gradient 2.25 -> 2.15 s; taking blockIdx.x % 8 as
// the XCD gave 6.02 and 2.16 s.)  `next` (the workgroup's next range to try,
// 0 = its own) lives in LDS across calls; `counter` holds kQueues zeroed
// ints.  Called by one lane of the workgroup; returns `total` once every
// range is empty.
constexpr int kQueues = 8;

__device__ inline int xcc_id() {
  int id;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(id));
  return id;
}

__device__ inline int pop_item(int* counter, int total, int& next) {
  const int xcd = xcc_id() & (kQueues - 1);
  for (; next < kQueues; ++next) {
    const int q = (xcd + next) & (kQueues - 1);
    const int start = (int)((long long)q * total / kQueues);
    const int len = (int)((long long)(q + 1) * total / kQueues) - start;
    if (len <= 0) continue;
    const int t = atomicAdd(counter + q, 1);
    if (t < len) return start + t;
  }
  return total;
}

}  // namespace pjz
