// Host build of the tiled search's (key, index) list words (cuda-knn-aniso-gaussian-feature-
// aggregation_amd/csrc/aknn_order.h, the same source the search kernel compiles) for the CPU
// test of the encoding.
#include "aknn_order.h"

extern "C" {

void order_words(const float* c, const int* i, long n, unsigned long long* out) {
  for (long j = 0; j < n; ++j) out[j] = aknn::list_word(c[j], i[j]);
}
void order_decode(const unsigned long long* w, long n, float* key, int* idx) {
  for (long j = 0; j < n; ++j) {
    key[j] = aknn::list_word_key(w[j]);
    idx[j] = aknn::list_word_index(w[j]);
  }
}
}
