// data_generation -- writes a synthetic checkpoint series generated on the
// GPU (the reference's src/data_generation.cpp over
// include/data_generation.hpp:13-95).
//
//   data_generation data_len num_chkpts mode num_changes prefix [--offsets-only]
//
// Checkpoint 0 is random bytes; checkpoint i > 0 perturbs checkpoint i-1 by
// `mode` with `num_changes` as its size parameter:
//   R  every byte random again
//   B  bytes [num_changes, data_len) random, the first num_changes kept
//   C  a window of num_changes bytes at a random offset random
//   I  nothing (identical)
//   S  num_changes random positions changed to a different byte value
//   Z  a window of num_changes bytes at a random offset zeroed
//   W  bytes [0, num_changes) swapped with a 128-aligned window at a random
//      offset past them
// and is written to "<prefix>.<i>.chkpt".
//
// Window offsets come from std::default_random_engine seeded with 1931 and
// one std::uniform_int_distribution<uint64_t>(0, data_len - num_changes)
// draw per checkpoint for C and Z (a rejection loop for W), the reference's
// calls in the reference's order, so a libstdc++ build picks the reference's
// offsets.  Byte values come from a counter-based generator (splitmix64 of
// seed, checkpoint and position), so a series is reproducible on any device;
// the reference draws them from a Kokkos random pool whose per-thread states
// make its values depend on scheduling.  --offsets-only prints the offsets
// and writes nothing (no GPU use).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr uint64_t kSeed = 1931;

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// splitmix64 finalizer (Steele, Lea, Flood: "Fast splittable pseudorandom
// number generators", OOPSLA 2014).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// 64 random bits for (stream, counter): stream = checkpoint index and purpose.
__host__ __device__ inline uint64_t rand64(uint64_t stream, uint64_t ctr) {
  return mix64(mix64(kSeed + stream * 0x9e3779b97f4a7c15ULL) + ctr * 0xd1b54a32d192ed03ULL);
}

// Bytes [begin, end) of p get byte (pos % 8) of rand64(stream, pos / 8): each
// lane writes one aligned 8-byte word, whole words with one store.
__global__ __launch_bounds__(256) void k_fill_random(uint8_t* __restrict__ p, uint64_t begin, uint64_t end,
                                                     uint64_t stream) {
  const uint64_t w0 = begin / 8, w1 = (end + 7) / 8;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w = w0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < w1; w += stride) {
    const uint64_t v = rand64(stream, w);
    const uint64_t a = w * 8;
    if (a >= begin && a + 8 <= end) {
      *reinterpret_cast<uint64_t*>(p + a) = v;
    } else {
      for (uint64_t b = 0; b < 8; b++)
        if (a + b >= begin && a + b < end) p[a + b] = (uint8_t)(v >> (8 * b));
    }
  }
}

// Sparse changes, two kernels so that repeated positions stay deterministic:
// the first only reads (position j and its new value, a function of the
// position and its old byte, never equal to it), the second writes.
__global__ __launch_bounds__(256) void k_sparse_pick(const uint8_t* __restrict__ p, uint64_t len, uint64_t n,
                                                     uint64_t stream, uint64_t* __restrict__ pos,
                                                     uint8_t* __restrict__ val) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const uint64_t q = rand64(stream, j) % len;
    const uint8_t old = p[q];
    pos[j] = q;
    val[j] = (uint8_t)(old + 1 + rand64(stream + 0x5eed, q) % 255);
  }
}

__global__ __launch_bounds__(256) void k_sparse_apply(uint8_t* __restrict__ p, uint64_t n,
                                                      const uint64_t* __restrict__ pos,
                                                      const uint8_t* __restrict__ val) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) p[pos[j]] = val[j];
}

uint32_t grid_for(uint64_t items) {
  const uint64_t g = (items + 255) / 256;
  return (uint32_t)(g == 0 ? 1 : g > 16384 ? 16384 : g);
}

void fill_random(uint8_t* p, uint64_t begin, uint64_t end, uint64_t stream, hipStream_t s) {
  if (end <= begin) return;
  hipLaunchKernelGGL(k_fill_random, dim3(grid_for((end - begin) / 8 + 2)), dim3(256), 0, s, p, begin, end, stream);
  hip_check(hipGetLastError(), "k_fill_random");
}

// The reference's offset draws (include/data_generation.hpp:50-52,60-62,85-91).
struct Offsets {
  std::default_random_engine gen{kSeed};
  uint64_t window(uint64_t len, uint64_t n) {
    std::uniform_int_distribution<uint64_t> d(0, len - n);
    return d(gen);
  }
  uint64_t swap_target(uint64_t len, uint64_t n) {
    std::uniform_int_distribution<uint64_t> d(0, len - n);
    uint64_t b = d(gen);
    while (b < n || b + n > len || b % 128 != 0) b = d(gen);
    return b;
  }
};

// A 128-aligned swap target at or past n with room for n bytes exists.
bool swap_possible(uint64_t len, uint64_t n) {
  const uint64_t first = (n + 127) / 128 * 128;
  return first + n <= len;
}

void write_file(const std::string& name, const uint8_t* host, uint64_t len) {
  FILE* f = fopen(name.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + name);
  const size_t w = fwrite(host, 1, len, f);
  fclose(f);
  if (w != len) throw std::runtime_error("short write to " + name);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s data_len num_chkpts mode(R|B|C|I|S|Z|W) num_changes prefix [--offsets-only]\n",
            argv[0]);
    return 2;
  }
  const uint64_t len = strtoull(argv[1], nullptr, 0);
  const uint32_t nckpt = (uint32_t)strtoul(argv[2], nullptr, 0);
  const char mode = argv[3][0];
  const uint64_t nc = strtoull(argv[4], nullptr, 0);
  const std::string prefix = std::string(argv[5]) + ".";
  const bool offsets_only = argc > 6 && !strcmp(argv[6], "--offsets-only");
  if (!strchr("RBCISZW", mode) || argv[3][1] != 0) {
    fprintf(stderr, "ERROR: mode must be one of R B C I S Z W\n");
    return 2;
  }
  if (len == 0 || nckpt == 0 || ((mode == 'B' || mode == 'C' || mode == 'Z' || mode == 'W') && nc > len) ||
      (mode == 'W' && !swap_possible(len, nc))) {
    fprintf(stderr, "ERROR: need data_len > 0, num_chkpts > 0 and num_changes that fits the mode\n");
    return 2;
  }
  printf("Data length: %lu\n", (unsigned long)len);
  printf("Number of checkpoints: %u\n", nckpt);
  printf("Mode: %c\n", mode);
  printf("Num changes: %lu\n", (unsigned long)nc);
  printf("File name: %s\n", prefix.c_str());
  Offsets off;
  if (offsets_only) {
    for (uint32_t i = 1; i < nckpt; i++) {
      if (mode == 'C' || mode == 'Z') printf("Offset: %lu\n", (unsigned long)off.window(len, nc));
      if (mode == 'W') printf("B offset: %lu\n", (unsigned long)off.swap_target(len, nc));
    }
    return 0;
  }
  try {
    hipStream_t s;
    hip_check(hipStreamCreate(&s), "hipStreamCreate");
    uint8_t *dev = nullptr, *host = nullptr, *tmp = nullptr;
    uint64_t* spos = nullptr;
    uint8_t* sval = nullptr;
    hip_check(hipMalloc(&dev, len), "hipMalloc");
    hip_check(hipHostMalloc((void**)&host, len, hipHostMallocDefault), "hipHostMalloc");
    if (mode == 'W' && nc) hip_check(hipMalloc(&tmp, nc), "hipMalloc");
    if (mode == 'S' && nc) {
      hip_check(hipMalloc(&spos, nc * 8), "hipMalloc");
      hip_check(hipMalloc(&sval, nc), "hipMalloc");
    }
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    double gen_ms = 0;
    for (uint32_t i = 0; i < nckpt; i++) {
      hip_check(hipEventRecord(e0, s), "hipEventRecord");
      const uint64_t stream = 2ull * i;  // one stream of random words per checkpoint
      if (i == 0 || mode == 'R') {
        fill_random(dev, 0, len, stream, s);
      } else if (mode == 'B') {
        fill_random(dev, nc, len, stream, s);
      } else if (mode == 'C') {
        const uint64_t o = off.window(len, nc);
        printf("Offset: %lu\n", (unsigned long)o);
        fill_random(dev, o, o + nc, stream, s);
      } else if (mode == 'Z') {
        const uint64_t o = off.window(len, nc);
        printf("Offset: %lu\n", (unsigned long)o);
        if (nc) hip_check(hipMemsetAsync(dev + o, 0, nc, s), "hipMemsetAsync");
      } else if (mode == 'S') {
        if (nc) {
          hipLaunchKernelGGL(k_sparse_pick, dim3(grid_for(nc)), dim3(256), 0, s, dev, len, nc, stream, spos, sval);
          hipLaunchKernelGGL(k_sparse_apply, dim3(grid_for(nc)), dim3(256), 0, s, dev, nc, spos, sval);
          hip_check(hipGetLastError(), "sparse kernels");
        }
      } else if (mode == 'W') {
        const uint64_t b = off.swap_target(len, nc);
        printf("B offset: %lu\n", (unsigned long)b);
        if (nc) {
          hip_check(hipMemcpyAsync(tmp, dev, nc, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
          hip_check(hipMemcpyAsync(dev, dev + b, nc, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
          hip_check(hipMemcpyAsync(dev + b, tmp, nc, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
        }
      }  // 'I': identical
      hip_check(hipEventRecord(e1, s), "hipEventRecord");
      hip_check(hipMemcpyAsync(host, dev, len, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      gen_ms += ms;
      write_file(prefix + std::to_string(i) + ".chkpt", host, len);
    }
    printf("Generated %u checkpoints of %lu bytes in %.3f ms of GPU time\n", nckpt, (unsigned long)len, gen_ms);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(dev);
    (void)hipFree(tmp);
    (void)hipFree(spos);
    (void)hipFree(sval);
    (void)hipHostFree(host);
    (void)hipStreamDestroy(s);
  } catch (const std::exception& e) {
    fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
  return 0;
}
