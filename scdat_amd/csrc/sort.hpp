// Launch interface of the cell sort.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace scdat {

template <typename T>
struct SoA {
  T *x, *y, *z, *vx, *vy, *vz;
  int32_t* cell;
  int32_t* id;
};

template <typename T>
struct SortArgs {
  SoA<T> src, dst;
  int64_t n;        // source slots
  int64_t ncell;    // tets
  uint32_t* count;  // [ncell]
  uint32_t* offset; // [ncell]
  uint32_t* cursor; // [ncell]
  uint32_t* tiles;  // [sort_tiles(ncell)]
  uint32_t* total;  // [1] alive count after the sort
};

template <typename T>
void launch_sort(const SortArgs<T>& a, hipStream_t s);
size_t sort_tiles(int64_t ncell);

}  // namespace scdat
