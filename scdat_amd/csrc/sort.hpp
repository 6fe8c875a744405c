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

// One particle as a record, for the sort's random-access pass: one 32-byte
// (FP32) or 64-byte (FP64) read per particle instead of eight scattered
// 4- or 8-byte reads.
template <typename T>
struct alignas(sizeof(T) == 4 ? 32 : 64) SortRec {
  T x, y, z, vx, vy, vz;
  int32_t cell, id;
};

template <typename T>
struct SortArgs {
  SoA<T> src, dst;
  int64_t n;        // source slots
  int64_t ncell;    // tets
  const int32_t* key;  // [ncell] rank of each cell in the sorted order (null: the cell index)
  uint32_t* count;  // [ncell]
  uint32_t* offset; // [ncell]
  uint32_t* cursor; // [ncell]
  uint32_t* tiles;  // [sort_tiles(ncell)]
  uint32_t* total;  // [1] alive count after the sort
  uint32_t* perm;   // [n] source slot of each destination slot
  SortRec<T>* stage;  // [n] the particles as records, in source order
  // captured steps: the source slot count on the device (n then only sizes
  // the grids); the alive count is stored back into it after the sort
  int64_t* n_dev;
};

template <typename T>
void launch_sort(const SortArgs<T>& a, hipStream_t s);
size_t sort_tiles(int64_t ncell);

}  // namespace scdat
