// Oracle part 4: the outer reorder's sort (sort.cpp:3-7): 1-based `index` ordered by
// descending arr[index-1] with std::sort, whose tie order is the libstdc++ introsort's.
// Test infrastructure only.
#include <algorithm>
#include "oracle.h"

extern "C" void orc_qsort(int* arr, int* index, int n) {
    std::sort(index, index + n, [arr](int a, int b) { return arr[a - 1] > arr[b - 1]; });
}
