// dd_device.hpp -- device-side building blocks of the tree-approach deduplicator
// for CDNA4 (gfx950): chunk hashes, heap-tree geometry and the digest map.
//
// Everything here is integer work.  Nothing is routed through MFMA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dd {

// Labels, include/utils.hpp:51-57.  PENDING marks a leaf whose digest must be
// inserted into the map before its label is known; FIXED|PENDING is a leaf of
// a baseline checkpoint whose digest equals the all-zero previous digest.
enum : uint8_t {
  FIRST_OCUR = 0,
  FIXED_DUPL = 1,
  SHIFT_DUPL = 2,
  FIRST_DUPL = 3,
  DONE = 4,
  PENDING = 0x80
};

enum : int { HASH_MURMUR3 = 0, HASH_MD5 = 1 };

// Map / pipeline status flags.
// FLAG_BAD_DIFF: a restart's parse found a diff inconsistent.
// FLAG_BULK_OFF: a baseline's bulk map build found a bucket with more leaves
// than its workgroup holds (raised with FLAG_MAP_FULL); the host runs the job
// again with the dense inserts.
// FLAG_BULK_DUPS: a bulk build without the leaf duplicate search found a
// repeated leaf digest (raised with FLAG_MAP_FULL, which stops the later
// stages); the host runs the job again with the search.
enum : uint32_t {
  FLAG_NEED_GROW = 1u,
  FLAG_MAP_FULL = 2u,
  FLAG_BAD_SOURCE = 4u,
  FLAG_BAD_DIFF = 8u,
  FLAG_BULK_OFF = 16u,
  FLAG_BULK_DUPS = 32u
};

struct Digest {
  uint64_t k0, k1;
};

// One open-addressing slot.  `val` packs NodeID{node, tree}
// (include/map_helpers.hpp:38-58) as tree<<32 | node, so a 64-bit atomicMin on
// it lowers the node id of an entry of the same checkpoint.
struct alignas(32) Slot {
  uint64_t k0, k1;
  unsigned long long val;
  uint64_t pad;
};

constexpr unsigned long long SLOT_EMPTY = ~0ull;

// Pointer kinds.  The host fills job tables and map views with plain
// pointers; the kernels read the same bytes with every pointer typed as a
// global-address-space pointer, so that loads, stores and atomics through
// them compile to global_* instructions.  A pointer read from memory is
// otherwise generic, and generic (flat) instructions also count against the
// LDS/scalar-memory counter: every wait for an LDS access, a shuffle, a job
// field or a barrier would then also wait for the loads and stores in flight.
template <class T>
using HostPtr = T*;
template <class T>
using GlobalPtr = __attribute__((address_space(1))) T*;

// Atomics through global pointers, with the semantics of HIP's atomicAdd and
// friends (relaxed order, device scope).
#define DD_GLOBAL_ATOMIC(name, builtin, T)                                       \
  __device__ __forceinline__ T name(GlobalPtr<T> p, T v) {                       \
    return builtin(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);            \
  }
DD_GLOBAL_ATOMIC(g_add, __hip_atomic_fetch_add, uint32_t)
DD_GLOBAL_ATOMIC(g_add, __hip_atomic_fetch_add, unsigned long long)
DD_GLOBAL_ATOMIC(g_or, __hip_atomic_fetch_or, uint32_t)
DD_GLOBAL_ATOMIC(g_min, __hip_atomic_fetch_min, unsigned long long)
DD_GLOBAL_ATOMIC(g_max, __hip_atomic_fetch_max, unsigned long long)
DD_GLOBAL_ATOMIC(g_exch, __hip_atomic_exchange, uint32_t)
#undef DD_GLOBAL_ATOMIC
// Compare-and-swap; returns the old value.
__device__ __forceinline__ unsigned long long g_cas(GlobalPtr<unsigned long long> p, unsigned long long cmp,
                                                    unsigned long long val) {
  __hip_atomic_compare_exchange_strong(p, &cmp, val, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return cmp;
}

// Streaming reads through the global address space (chunk data pointers are
// generic where they come from byte-range helpers).
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4_t global_u32x4;
// 16 bytes at p (16-byte aligned) through a global load.
__device__ __forceinline__ uint4 gload16(const void* p) {
  const u32x4_t v = *(const global_u32x4*)(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gload16(GlobalPtr<const void> p) {
  const u32x4_t v = *(const global_u32x4*)(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ Digest load_digest(GlobalPtr<const Digest> p) {
  const uint4 v = gload16((GlobalPtr<const void>)p);
  return Digest{v.x | ((uint64_t)v.y << 32), v.z | ((uint64_t)v.w << 32)};
}
// One 16-byte store (digest arrays are 16-byte aligned).
__device__ __forceinline__ void store_digest(GlobalPtr<Digest> p, const Digest& d) {
  *(GlobalPtr<u32x4_t>)p = u32x4_t{(uint32_t)d.k0, (uint32_t)(d.k0 >> 32), (uint32_t)d.k1, (uint32_t)(d.k1 >> 32)};
}

__host__ __device__ inline unsigned long long pack_node(uint32_t tree, uint32_t node) {
  return ((unsigned long long)tree << 32) | node;
}

// ---------------------------------------------------------------------------
// Murmur3 x64-128 with the reference's truncating rotate
// (include/kokkos_murmur3.hpp:30-31: rotl64 returns uint32_t).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t rotq(uint64_t x, int r) {
  return (uint64_t)(uint32_t)((x << r) | (x >> (64 - r)));
}

__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

struct Murmur {
  uint64_t h1 = 0, h2 = 0;
  static constexpr uint64_t C1 = 0x87c37b91114253d5ull;
  static constexpr uint64_t C2 = 0x4cf5ad432745937full;

  __device__ __forceinline__ void block(uint64_t k1, uint64_t k2) {
    k1 *= C1; k1 = rotq(k1, 31); k1 *= C2; h1 ^= k1;
    h1 = rotq(h1, 27); h1 += h2; h1 = h1 * 5 + 0x52dce729;
    k2 *= C2; k2 = rotq(k2, 33); k2 *= C1; h2 ^= k2;
    h2 = rotq(h2, 31); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
  }
  // Tail of `rem` (< 16) bytes given as two little-endian words (bytes past
  // `rem` must be zero).
  __device__ __forceinline__ void tail(uint64_t t0, uint64_t t1, uint32_t rem) {
    if (rem > 8) {
      uint64_t k2 = t1;
      k2 *= C2; k2 = rotq(k2, 33); k2 *= C1; h2 ^= k2;
    }
    if (rem > 0) {
      uint64_t k1 = t0;
      k1 *= C1; k1 = rotq(k1, 31); k1 *= C2; h1 ^= k1;
    }
  }
  __device__ __forceinline__ Digest finish(uint64_t len) {
    h1 ^= len; h2 ^= len;
    h1 += h2; h2 += h1;
    h1 = fmix64(h1); h2 = fmix64(h2);
    h1 += h2; h2 += h1;
    return Digest{h1, h2};
  }
};

// ---------------------------------------------------------------------------
// MD5 (RFC 1321).  The reference's version (include/kokkos_md5.hpp:60-265)
// equals it for every length below INT_MAX, which covers chunks and nodes.
// ---------------------------------------------------------------------------
#define DD_MD5_STEP(f, a, b, c, d, x, t, s) \
  a += f(b, c, d) + (x) + (t);              \
  a = (a << (s)) | (a >> (32 - (s)));       \
  a += b;
#define DD_F1(x, y, z) ((z) ^ ((x) & ((y) ^ (z))))
#define DD_F2(x, y, z) ((y) ^ ((z) & ((x) ^ (y))))
#define DD_F3(x, y, z) ((x) ^ (y) ^ (z))
#define DD_F4(x, y, z) ((y) ^ ((x) | ~(z)))

__device__ __forceinline__ void md5_compress(uint32_t st[4], const uint32_t X[16]) {
  uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
  DD_MD5_STEP(DD_F1, a, b, c, d, X[0], 0xd76aa478, 7)
  DD_MD5_STEP(DD_F1, d, a, b, c, X[1], 0xe8c7b756, 12)
  DD_MD5_STEP(DD_F1, c, d, a, b, X[2], 0x242070db, 17)
  DD_MD5_STEP(DD_F1, b, c, d, a, X[3], 0xc1bdceee, 22)
  DD_MD5_STEP(DD_F1, a, b, c, d, X[4], 0xf57c0faf, 7)
  DD_MD5_STEP(DD_F1, d, a, b, c, X[5], 0x4787c62a, 12)
  DD_MD5_STEP(DD_F1, c, d, a, b, X[6], 0xa8304613, 17)
  DD_MD5_STEP(DD_F1, b, c, d, a, X[7], 0xfd469501, 22)
  DD_MD5_STEP(DD_F1, a, b, c, d, X[8], 0x698098d8, 7)
  DD_MD5_STEP(DD_F1, d, a, b, c, X[9], 0x8b44f7af, 12)
  DD_MD5_STEP(DD_F1, c, d, a, b, X[10], 0xffff5bb1, 17)
  DD_MD5_STEP(DD_F1, b, c, d, a, X[11], 0x895cd7be, 22)
  DD_MD5_STEP(DD_F1, a, b, c, d, X[12], 0x6b901122, 7)
  DD_MD5_STEP(DD_F1, d, a, b, c, X[13], 0xfd987193, 12)
  DD_MD5_STEP(DD_F1, c, d, a, b, X[14], 0xa679438e, 17)
  DD_MD5_STEP(DD_F1, b, c, d, a, X[15], 0x49b40821, 22)
  DD_MD5_STEP(DD_F2, a, b, c, d, X[1], 0xf61e2562, 5)
  DD_MD5_STEP(DD_F2, d, a, b, c, X[6], 0xc040b340, 9)
  DD_MD5_STEP(DD_F2, c, d, a, b, X[11], 0x265e5a51, 14)
  DD_MD5_STEP(DD_F2, b, c, d, a, X[0], 0xe9b6c7aa, 20)
  DD_MD5_STEP(DD_F2, a, b, c, d, X[5], 0xd62f105d, 5)
  DD_MD5_STEP(DD_F2, d, a, b, c, X[10], 0x02441453, 9)
  DD_MD5_STEP(DD_F2, c, d, a, b, X[15], 0xd8a1e681, 14)
  DD_MD5_STEP(DD_F2, b, c, d, a, X[4], 0xe7d3fbc8, 20)
  DD_MD5_STEP(DD_F2, a, b, c, d, X[9], 0x21e1cde6, 5)
  DD_MD5_STEP(DD_F2, d, a, b, c, X[14], 0xc33707d6, 9)
  DD_MD5_STEP(DD_F2, c, d, a, b, X[3], 0xf4d50d87, 14)
  DD_MD5_STEP(DD_F2, b, c, d, a, X[8], 0x455a14ed, 20)
  DD_MD5_STEP(DD_F2, a, b, c, d, X[13], 0xa9e3e905, 5)
  DD_MD5_STEP(DD_F2, d, a, b, c, X[2], 0xfcefa3f8, 9)
  DD_MD5_STEP(DD_F2, c, d, a, b, X[7], 0x676f02d9, 14)
  DD_MD5_STEP(DD_F2, b, c, d, a, X[12], 0x8d2a4c8a, 20)
  DD_MD5_STEP(DD_F3, a, b, c, d, X[5], 0xfffa3942, 4)
  DD_MD5_STEP(DD_F3, d, a, b, c, X[8], 0x8771f681, 11)
  DD_MD5_STEP(DD_F3, c, d, a, b, X[11], 0x6d9d6122, 16)
  DD_MD5_STEP(DD_F3, b, c, d, a, X[14], 0xfde5380c, 23)
  DD_MD5_STEP(DD_F3, a, b, c, d, X[1], 0xa4beea44, 4)
  DD_MD5_STEP(DD_F3, d, a, b, c, X[4], 0x4bdecfa9, 11)
  DD_MD5_STEP(DD_F3, c, d, a, b, X[7], 0xf6bb4b60, 16)
  DD_MD5_STEP(DD_F3, b, c, d, a, X[10], 0xbebfbc70, 23)
  DD_MD5_STEP(DD_F3, a, b, c, d, X[13], 0x289b7ec6, 4)
  DD_MD5_STEP(DD_F3, d, a, b, c, X[0], 0xeaa127fa, 11)
  DD_MD5_STEP(DD_F3, c, d, a, b, X[3], 0xd4ef3085, 16)
  DD_MD5_STEP(DD_F3, b, c, d, a, X[6], 0x04881d05, 23)
  DD_MD5_STEP(DD_F3, a, b, c, d, X[9], 0xd9d4d039, 4)
  DD_MD5_STEP(DD_F3, d, a, b, c, X[12], 0xe6db99e5, 11)
  DD_MD5_STEP(DD_F3, c, d, a, b, X[15], 0x1fa27cf8, 16)
  DD_MD5_STEP(DD_F3, b, c, d, a, X[2], 0xc4ac5665, 23)
  DD_MD5_STEP(DD_F4, a, b, c, d, X[0], 0xf4292244, 6)
  DD_MD5_STEP(DD_F4, d, a, b, c, X[7], 0x432aff97, 10)
  DD_MD5_STEP(DD_F4, c, d, a, b, X[14], 0xab9423a7, 15)
  DD_MD5_STEP(DD_F4, b, c, d, a, X[5], 0xfc93a039, 21)
  DD_MD5_STEP(DD_F4, a, b, c, d, X[12], 0x655b59c3, 6)
  DD_MD5_STEP(DD_F4, d, a, b, c, X[3], 0x8f0ccc92, 10)
  DD_MD5_STEP(DD_F4, c, d, a, b, X[10], 0xffeff47d, 15)
  DD_MD5_STEP(DD_F4, b, c, d, a, X[1], 0x85845dd1, 21)
  DD_MD5_STEP(DD_F4, a, b, c, d, X[8], 0x6fa87e4f, 6)
  DD_MD5_STEP(DD_F4, d, a, b, c, X[15], 0xfe2ce6e0, 10)
  DD_MD5_STEP(DD_F4, c, d, a, b, X[6], 0xa3014314, 15)
  DD_MD5_STEP(DD_F4, b, c, d, a, X[13], 0x4e0811a1, 21)
  DD_MD5_STEP(DD_F4, a, b, c, d, X[4], 0xf7537e82, 6)
  DD_MD5_STEP(DD_F4, d, a, b, c, X[11], 0xbd3af235, 10)
  DD_MD5_STEP(DD_F4, c, d, a, b, X[2], 0x2ad7d2bb, 15)
  DD_MD5_STEP(DD_F4, b, c, d, a, X[9], 0xeb86d391, 21)
  st[0] += a; st[1] += b; st[2] += c; st[3] += d;
}

__device__ __forceinline__ void md5_init(uint32_t st[4]) {
  st[0] = 0x67452301u; st[1] = 0xefcdab89u; st[2] = 0x98badcfeu; st[3] = 0x10325476u;
}

__device__ __forceinline__ Digest md5_digest(const uint32_t st[4]) {
  return Digest{(uint64_t)st[0] | ((uint64_t)st[1] << 32), (uint64_t)st[2] | ((uint64_t)st[3] << 32)};
}

// ---------------------------------------------------------------------------
// Byte-granular loads for chunks that are not 16-byte aligned.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t load_le64_bytes(const uint8_t* p, uint32_t n) {
  uint64_t v = 0;
  for (uint32_t i = 0; i < n; i++) v |= (uint64_t)p[i] << (8 * i);
  return v;
}

__device__ __forceinline__ uint32_t load_le32_bytes(const uint8_t* p, uint32_t n) {
  uint32_t v = 0;
  for (uint32_t i = 0; i < n; i++) v |= (uint32_t)p[i] << (8 * i);
  return v;
}

// Murmur3 of an arbitrary byte range (any alignment).
__device__ __forceinline__ Digest murmur3_bytes(const uint8_t* p, uint64_t len) {
  Murmur m;
  uint64_t nb = len / 16;
  for (uint64_t i = 0; i < nb; i++) m.block(load_le64_bytes(p + 16 * i, 8), load_le64_bytes(p + 16 * i + 8, 8));
  uint32_t rem = (uint32_t)(len & 15);
  const uint8_t* t = p + 16 * nb;
  uint64_t t0 = load_le64_bytes(t, rem < 8 ? rem : 8);
  uint64_t t1 = rem > 8 ? load_le64_bytes(t + 8, rem - 8) : 0;
  m.tail(t0, t1, rem);
  return m.finish(len);
}

// Murmur3 of 16-byte aligned data.
__device__ __forceinline__ Digest murmur3_aligned(const uint8_t* p, uint64_t len) {
  Murmur m;
  const uint8_t* q = p;  // 16-byte vectors through global loads
  uint64_t nb = len / 16;
  uint64_t i = 0;
  for (; i + 4 <= nb; i += 4) {
    uint4 a = gload16(q + 16 * i), b = gload16(q + 16 * (i + 1)), c = gload16(q + 16 * (i + 2)),
          d = gload16(q + 16 * (i + 3));
    m.block(a.x | ((uint64_t)a.y << 32), a.z | ((uint64_t)a.w << 32));
    m.block(b.x | ((uint64_t)b.y << 32), b.z | ((uint64_t)b.w << 32));
    m.block(c.x | ((uint64_t)c.y << 32), c.z | ((uint64_t)c.w << 32));
    m.block(d.x | ((uint64_t)d.y << 32), d.z | ((uint64_t)d.w << 32));
  }
  for (; i < nb; i++) {
    uint4 a = gload16(q + 16 * i);
    m.block(a.x | ((uint64_t)a.y << 32), a.z | ((uint64_t)a.w << 32));
  }
  uint32_t rem = (uint32_t)(len & 15);
  if (rem) {
    const uint8_t* t = p + 16 * nb;
    uint64_t t0 = load_le64_bytes(t, rem < 8 ? rem : 8);
    uint64_t t1 = rem > 8 ? load_le64_bytes(t + 8, rem - 8) : 0;
    m.tail(t0, t1, rem);
  }
  return m.finish(len);
}

// Murmur3 of 16-byte aligned data with the loads of the next 128 bytes issued
// before the current 128 bytes are hashed (double buffering): for long chunks
// one lane then keeps 256 bytes in flight instead of 64.
__device__ __forceinline__ Digest murmur3_aligned_deep(const uint8_t* p, uint64_t len) {
  Murmur m;
  const uint8_t* q = p;  // 16-byte vectors through global loads
  const uint64_t nb = len / 16;
  uint64_t i = 0;
  if (nb >= 8) {
    uint4 cur[8];
#pragma unroll
    for (int k = 0; k < 8; k++) cur[k] = gload16(q + 16 * k);
    for (; i + 16 <= nb; i += 8) {
      uint4 nxt[8];
#pragma unroll
      for (int k = 0; k < 8; k++) nxt[k] = gload16(q + 16 * (i + 8 + k));
#pragma unroll
      for (int k = 0; k < 8; k++) m.block(cur[k].x | ((uint64_t)cur[k].y << 32), cur[k].z | ((uint64_t)cur[k].w << 32));
#pragma unroll
      for (int k = 0; k < 8; k++) cur[k] = nxt[k];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) m.block(cur[k].x | ((uint64_t)cur[k].y << 32), cur[k].z | ((uint64_t)cur[k].w << 32));
    i += 8;
  }
  for (; i < nb; i++) {
    const uint4 a = gload16(q + 16 * i);
    m.block(a.x | ((uint64_t)a.y << 32), a.z | ((uint64_t)a.w << 32));
  }
  const uint32_t rem = (uint32_t)(len & 15);
  if (rem) {
    const uint8_t* t = p + 16 * nb;
    const uint64_t t0 = load_le64_bytes(t, rem < 8 ? rem : 8);
    const uint64_t t1 = rem > 8 ? load_le64_bytes(t + 8, rem - 8) : 0;
    m.tail(t0, t1, rem);
  }
  return m.finish(len);
}

// Murmur3 of a whole chunk of 16·n bytes (n <= 8, 16-byte aligned): every
// load is issued before any hashing, as in the deep variant, but without its
// second buffer, so the kernel needs fewer registers and keeps more waves per
// SIMD (chunks of at most 128 bytes).
// Straight-line code per length: a load behind a per-vector condition would
// be issued after the previous one had arrived.
template <int N>
__device__ __forceinline__ Digest murmur3_fixed(const uint8_t* p) {
  uint4 v[N];
#pragma unroll
  for (int k = 0; k < N; k++) v[k] = gload16(p + 16 * k);
  Murmur m;
#pragma unroll
  for (int k = 0; k < N; k++) m.block(v[k].x | ((uint64_t)v[k].y << 32), v[k].z | ((uint64_t)v[k].w << 32));
  return m.finish(16ull * N);
}

__device__ __forceinline__ Digest murmur3_short(const uint8_t* p, uint32_t n) {
  switch (n) {
    case 8: return murmur3_fixed<8>(p);
    case 7: return murmur3_fixed<7>(p);
    case 6: return murmur3_fixed<6>(p);
    case 5: return murmur3_fixed<5>(p);
    case 4: return murmur3_fixed<4>(p);
    case 3: return murmur3_fixed<3>(p);
    case 2: return murmur3_fixed<2>(p);
    default: return murmur3_fixed<1>(p);
  }
}

// MD5 of an arbitrary byte range; `Aligned` promises 16-byte alignment.
template <bool Aligned>
__device__ __forceinline__ Digest md5_range(const uint8_t* p, uint64_t len) {
  uint32_t st[4];
  md5_init(st);
  uint64_t nfull = len / 64;
  uint32_t X[16];
  for (uint64_t b = 0; b < nfull; b++) {
    const uint8_t* q = p + 64 * b;
    if (Aligned) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        uint4 w = gload16(q + 16 * i);
        X[4 * i] = w.x; X[4 * i + 1] = w.y; X[4 * i + 2] = w.z; X[4 * i + 3] = w.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; i++) X[i] = load_le32_bytes(q + 4 * i, 4);
    }
    md5_compress(st, X);
  }
  // Final block(s): tail bytes, 0x80, zeros, 64-bit bit length.
  uint32_t rem = (uint32_t)(len - 64 * nfull);
  const uint8_t* t = p + 64 * nfull;
  uint64_t bits = len * 8;
  uint32_t nwords = rem / 4, rb = rem & 3;
  uint32_t partial = rb ? load_le32_bytes(t + 4 * nwords, rb) : 0;
  partial |= 0x80u << (8 * rb);
#pragma unroll
  for (int i = 0; i < 16; i++) {
    uint32_t w = 0;
    if ((uint32_t)i < nwords) w = Aligned ? reinterpret_cast<const uint32_t*>(t)[i] : load_le32_bytes(t + 4 * i, 4);
    else if ((uint32_t)i == nwords) w = partial;
    X[i] = w;
  }
  if (rem < 56) {
    X[14] = (uint32_t)bits;
    X[15] = (uint32_t)(bits >> 32);
    md5_compress(st, X);
  } else {
    md5_compress(st, X);
#pragma unroll
    for (int i = 0; i < 14; i++) X[i] = 0;
    X[14] = (uint32_t)bits;
    X[15] = (uint32_t)(bits >> 32);
    md5_compress(st, X);
  }
  return md5_digest(st);
}

// Digest of an interior node: hash of the two child digests (32 bytes), as
// hash((uint8_t*)&tree(child_l), 2*sizeof(HashDigest), ...) in
// include/reference_impl.hpp:79,96.
__device__ __forceinline__ Digest hash_pair(int kind, const Digest& l, const Digest& r) {
  if (kind == HASH_MD5) {
    uint32_t st[4];
    md5_init(st);
    uint32_t X[16] = {(uint32_t)l.k0, (uint32_t)(l.k0 >> 32), (uint32_t)l.k1, (uint32_t)(l.k1 >> 32),
                      (uint32_t)r.k0, (uint32_t)(r.k0 >> 32), (uint32_t)r.k1, (uint32_t)(r.k1 >> 32),
                      0x80u, 0, 0, 0, 0, 0, 256u, 0};
    md5_compress(st, X);
    return md5_digest(st);
  }
  Murmur m;
  m.block(l.k0, l.k1);
  m.block(r.k0, r.k1);
  return m.finish(32);
}

__device__ __forceinline__ bool digest_eq(const Digest& a, const Digest& b) {
  return a.k0 == b.k0 && a.k1 == b.k1;
}

// ---------------------------------------------------------------------------
// Heap-ordered tree of 2C-1 nodes: leaf of chunk k is node C-1+k, children of
// node i are 2i+1 and 2i+2 (include/kokkos_merkle_tree.hpp:16-31).
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t ilog2_64(uint64_t x) {
#ifdef __HIP_DEVICE_COMPILE__
  return 63 - __clzll((long long)x);
#else
  return 63 - __builtin_clzll(x);
#endif
}

struct TreeGeom {
  uint64_t C, N;  // chunks, nodes
  uint32_t D;     // depth of the deepest level
};

__host__ __device__ inline TreeGeom make_geom(uint64_t C) {
  TreeGeom g;
  g.C = C;
  g.N = 2 * C - 1;
  g.D = ilog2_64(g.N);
  return g;
}

// Leaf runs of a subtree in left-to-right order: `n0` nodes starting at `a0`
// on the deepest level, then `n1` nodes starting at `a1` one level up.
struct LeafRuns {
  uint64_t a0, n0, a1, n1;
};

// Rank of chunk k among the root's leaves in left-to-right order (its place
// in the data section of a diff whose only region is the root).  The deepest
// level holds chunks kd..C-1 (kd = 2^D - C) and comes first, then the level
// above with chunks 0..kd-1: subtree_runs(g, 0) as a rotation.
__host__ __device__ inline uint64_t root_rank(const TreeGeom& g, uint64_t k) {
  const uint64_t kd = (1ull << g.D) - g.C;
  return k >= kd ? k - kd : k + (g.C - kd);
}

__host__ __device__ inline LeafRuns subtree_runs(const TreeGeom& g, uint64_t v) {
  LeafRuns r{0, 0, 0, 0};
  if (v >= g.C - 1) {  // a leaf
    r.a0 = v;
    r.n0 = 1;
    return r;
  }
  uint32_t d = ilog2_64(v + 1);
  uint64_t lo = ((v + 1) << (g.D - d)) - 1;
  uint64_t hi = ((v + 2) << (g.D - d)) - 2;
  if (hi > g.N - 1) hi = g.N - 1;
  if (lo <= g.N - 1) {
    r.a0 = lo;
    r.n0 = hi - lo + 1;
  }
  if (d + 1 <= g.D) {
    uint64_t lo1 = ((v + 1) << (g.D - 1 - d)) - 1;
    uint64_t hi1 = ((v + 2) << (g.D - 1 - d)) - 2;
    if (lo1 < g.C - 1) lo1 = g.C - 1;
    if (lo1 <= hi1) {
      r.a1 = lo1;
      r.n1 = hi1 - lo1 + 1;
    }
  }
  return r;
}

__host__ __device__ inline uint64_t run_leaf(const LeafRuns& r, uint64_t idx) {
  return idx < r.n0 ? r.a0 + idx : r.a1 + (idx - r.n0);
}

// ---------------------------------------------------------------------------
// Digest map probing.
// ---------------------------------------------------------------------------
// The map is split into buckets of kBucketSlots consecutive slots (the whole
// map when it is smaller): a probe sequence starts at the home slot and wraps
// around inside the home slot's bucket, so each bucket is an independent
// linear-probing table.  A baseline builds every bucket at once in LDS and
// writes it out whole (k_bulk_build).  The capacity is a multiple of
// kBucketSlots (a power of two of 1024 .. 4096 slots for small maps), not a
// power of two: the bucket of a digest comes from a multiply-high of its
// mixed key by the bucket count.
constexpr uint32_t kBucketLog = 12;
constexpr uint32_t kBucketSlots = 1u << kBucketLog;

template <template <class> class P>
struct MapT {
  P<Slot> slots;
  uint64_t nbuckets;  // capacity / (bmask + 1)
  uint64_t bmask;     // slots per bucket - 1
};
using MapView = MapT<HostPtr>;  // as the host fills it
using MapG = MapT<GlobalPtr>;   // as the kernels read it
static_assert(sizeof(MapView) == sizeof(MapG), "one layout for both pointer kinds");

// Home slot of a digest: bucket from the high bits of the mixed key, slot in
// the bucket from its bits 20 and up.
template <class M>
__host__ __device__ inline uint64_t map_start(const M& m, const Digest& k) {
  const uint64_t h = (k.k0 ^ (k.k1 * 0x9e3779b97f4a7c15ull)) * 0xff51afd7ed558ccdull;
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t b = __umul64hi(h, m.nbuckets);
#else
  const uint64_t b = (uint64_t)(((unsigned __int128)h * m.nbuckets) >> 64);
#endif
  return b * (m.bmask + 1) + ((h >> 20) & m.bmask);
}

// Slot e places after s in s's bucket (wrapping inside the bucket).
template <class M>
__host__ __device__ inline uint64_t map_next(const M& m, uint64_t s, uint64_t e) {
  return (s & ~m.bmask) | ((s + e) & m.bmask);
}

}  // namespace dd
