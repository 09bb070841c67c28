// qsort_ (sort.cpp:3-7): order the 1-based `index` by descending arr(index).  The outer
// reorder's tie order is that of the C++ library's std::sort, so the same call with the
// same comparator is used; any other sort would change which particle goes where.
#include <algorithm>

#include "../../include/smcf.h"

extern "C" void qsort_(int* arr, int* index, int* n) {
  std::sort(index, index + *n, [arr](int a, int b) { return arr[a - 1] > arr[b - 1]; });
}
