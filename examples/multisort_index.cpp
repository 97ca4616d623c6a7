// MultiSort with index tracking through the drop-in header (the reference
// lists it as unsupported, README.md:46): double keys with duplicates,
// 2038-key arrays; every array must equal std::stable_sort of its
// (key, index) pairs by key.
#include <algorithm>
#include <cstdlib>
#include <iostream>
#include <utility>
#include <vector>
#include "turtle-sort.hpp"

int main(int argc, char **argv)
{
    using Type = double;
    constexpr int arrSize = 2038;
    constexpr int blockSize = 32;
    const int numArrays = argc > 1 ? std::atoi(argv[1]) : 4000;
    const int n = arrSize * numArrays;
    Turtle::TurtleSort<Type, true> sorter(n, false);
    std::vector<Type> keys(n);
    std::vector<int> index(n);
    srand(777);
    for (int i = 0; i < n; i++) {
        keys[i] = (double)(rand() % 1000) * 0.25 - 100.0;   // many equal keys
        index[i] = i;
    }
    std::vector<std::pair<Type, int>> expect(n);
    for (int i = 0; i < n; i++) expect[i] = {keys[i], index[i]};
    for (int a = 0; a < numArrays; a++)
        std::stable_sort(expect.begin() + (size_t)a * arrSize, expect.begin() + (size_t)(a + 1) * arrSize,
                         [](const std::pair<Type, int> &x, const std::pair<Type, int> &y) { return x.first < y.first; });
    double seconds = sorter.MultiSort<Type, arrSize, blockSize>(numArrays, keys.data(), index.data());
    std::cout << "Sorting " << numArrays << " arrays of " << arrSize << " elements with index took " << seconds
              << " seconds" << std::endl;
    for (int i = 0; i < n; i++) {
        if (keys[i] != expect[i].first || index[i] != expect[i].second) {
            std::cout << "sort failed at " << i << std::endl;
            return 1;
        }
    }
    std::cout << "sort success" << std::endl;
    return 0;
}
