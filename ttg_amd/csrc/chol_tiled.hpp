// Lower tile-packed layout shared by the tiled drivers and generators.
#pragma once

#include <hip/hip_runtime.h>

namespace ttgc {

// tile (m, n), n <= m, of an NT x NT tile grid; column-major over tiles.
__host__ __device__ inline long tile_index(int m, int n, int nt) {
  return (long)n * nt - (long)n * (n - 1) / 2 + (m - n);
}

hipError_t launch_prepare(hipStream_t s, int N, int NB, double *tiles, int *d_info);

}  // namespace ttgc
