// Top-K scan variants for float, key AKNN_KEY_EUCLIDEAN.
#include "aknn_knn_impl.h"

namespace aknn {
template hipError_t launch_knn_typed<float, AKNN_KEY_EUCLIDEAN>(const KnnLaunch&, hipStream_t);
}  // namespace aknn
