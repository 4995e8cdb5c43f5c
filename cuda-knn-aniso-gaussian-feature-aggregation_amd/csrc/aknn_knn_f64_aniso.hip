// Top-K scan variants for double, key AKNN_KEY_ANISO.
#include "aknn_knn_impl.h"

namespace aknn {
template hipError_t launch_knn_typed<double, AKNN_KEY_ANISO>(const KnnLaunch&, hipStream_t);
}  // namespace aknn
