// Dispatcher of the top-K scan and the point-record packing kernel.
#include "aknn_common.h"
#include "aknn_internal.h"

namespace aknn {

template <typename T, int KEY>
hipError_t launch_knn_typed(const KnnLaunch& L, hipStream_t st);

// ------------------------------------------------------------------ packing
template <typename T>
__global__ void pack_points_kernel(const T* __restrict__ p, const T* __restrict__ sigma,
                                   const T* __restrict__ R, T* __restrict__ rec, int64_t rows) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows * 16) return;
  const int64_t n = i >> 4;
  const int c = static_cast<int>(i & 15);
  T v;
  if (c < 3) v = p[n * 3 + c];
  else if (c < 12) v = R ? R[n * 9 + (c - 3)] : T(0);           // R / sigma may be absent for the
  else if (c < 15) v = sigma ? sigma[n * 3 + (c - 12)] : T(0);  // Euclidean key (never read then)
  else v = T(0);
  rec[i] = v;
}

template <typename T>
hipError_t launch_pack_points(const void* p, const void* sigma, const void* R, void* rec,
                              int64_t rows, hipStream_t st) {
  if (rows == 0) return hipSuccess;
  const int64_t n = rows * 16;
  const int bs = 256;
  hipLaunchKernelGGL((pack_points_kernel<T>), dim3(static_cast<unsigned>((n + bs - 1) / bs)), dim3(bs), 0,
                     st, static_cast<const T*>(p), static_cast<const T*>(sigma),
                     static_cast<const T*>(R), static_cast<T*>(rec), rows);
  return hipGetLastError();
}

template hipError_t launch_pack_points<float>(const void*, const void*, const void*, void*, int64_t, hipStream_t);
template hipError_t launch_pack_points<double>(const void*, const void*, const void*, void*, int64_t, hipStream_t);

// The forward's cache in one launch: thread i writes query order entry i, point order
// entry i (each copied from the search's order, or the identity when there is none) and
// point record value i (as pack_points_kernel); any of the three outputs may be absent.
template <typename T>
__global__ void cache_fill_kernel(const uint32_t* __restrict__ qord, int32_t* __restrict__ qout, int64_t qrows,
                                  const uint32_t* __restrict__ pord, int32_t* __restrict__ pout, int64_t prows,
                                  const T* __restrict__ p, const T* __restrict__ sigma, const T* __restrict__ R,
                                  T* __restrict__ rec) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (qout && i < qrows) qout[i] = qord ? static_cast<int32_t>(qord[i]) : static_cast<int32_t>(i);
  if (pout && i < prows) pout[i] = pord ? static_cast<int32_t>(pord[i]) : static_cast<int32_t>(i);
  if (rec && i < prows * 16) {
    const int64_t n = i >> 4;
    const int c = static_cast<int>(i & 15);
    T v;
    if (c < 3) v = p[n * 3 + c];
    else if (c < 12) v = R ? R[n * 9 + (c - 3)] : T(0);
    else if (c < 15) v = sigma ? sigma[n * 3 + (c - 12)] : T(0);
    else v = T(0);
    rec[i] = v;
  }
}

template <typename T>
hipError_t launch_cache_fill(const uint32_t* qord, int32_t* qout, int64_t qrows, const uint32_t* pord,
                             int32_t* pout, int64_t prows, const void* p, const void* sigma, const void* R,
                             void* rec, hipStream_t st) {
  int64_t n = 0;
  if (qout) n = qrows;
  if (pout && prows > n) n = prows;
  if (rec && prows * 16 > n) n = prows * 16;
  if (n == 0) return hipSuccess;
  const int bs = 256;
  hipLaunchKernelGGL((cache_fill_kernel<T>), dim3(static_cast<unsigned>((n + bs - 1) / bs)), dim3(bs), 0, st, qord,
                     qout, qrows, pord, pout, prows, static_cast<const T*>(p), static_cast<const T*>(sigma),
                     static_cast<const T*>(R), static_cast<T*>(rec));
  return hipGetLastError();
}

template hipError_t launch_cache_fill<float>(const uint32_t*, int32_t*, int64_t, const uint32_t*, int32_t*, int64_t,
                                             const void*, const void*, const void*, void*, hipStream_t);
template hipError_t launch_cache_fill<double>(const uint32_t*, int32_t*, int64_t, const uint32_t*, int32_t*, int64_t,
                                              const void*, const void*, const void*, void*, hipStream_t);

hipError_t launch_knn(const KnnLaunch& L, hipStream_t st) {
  if (L.dtype == AKNN_F32) {
    return L.key == AKNN_KEY_EUCLIDEAN ? launch_knn_typed<float, AKNN_KEY_EUCLIDEAN>(L, st)
                                       : launch_knn_typed<float, AKNN_KEY_ANISO>(L, st);
  }
  return L.key == AKNN_KEY_EUCLIDEAN ? launch_knn_typed<double, AKNN_KEY_EUCLIDEAN>(L, st)
                                     : launch_knn_typed<double, AKNN_KEY_ANISO>(L, st);
}

}  // namespace aknn
