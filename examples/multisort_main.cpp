// The reference's MultiSort test driver (main.cpp:4-55) against the drop-in
// header: 32,000 arrays x 2038 float keys, BlockSize 32, 10 repetitions, the
// same rand() * 3.14f fill.  Each array is checked for equality with
// std::sort of its input, not only for sortedness.
#include <algorithm>
#include <cstdlib>
#include <iostream>
#include <vector>
#include "turtle-sort.hpp"

int main(int argc, char **argv)
{
    using Type = float;
    constexpr int arrSize = 2038;
    constexpr int blockSize = 32;
    int numArraysToSort = 1000 * blockSize;
    int reps = argc > 1 ? std::atoi(argv[1]) : 10;
    int n = arrSize * numArraysToSort;
    bool compress = false;
    Turtle::TurtleSort<Type> sorter(n, compress);
    std::vector<Type> hostData(n), expect(n);
    srand(12345);
    for (int k = 0; k < reps; k++) {
        for (int i = 0; i < n; i++) hostData[i] = rand() * 3.14f;
        expect = hostData;
        for (int a = 0; a < numArraysToSort; a++)
            std::sort(expect.begin() + (size_t)a * arrSize, expect.begin() + (size_t)(a + 1) * arrSize);
        double seconds = sorter.MultiSort<Type, arrSize, blockSize>(numArraysToSort, hostData.data());
        std::cout << "Sorting " << numArraysToSort << " arrays of " << arrSize << " elements took " << seconds
                  << " seconds" << std::endl;
        if (hostData != expect) {
            std::cout << "sort failed" << std::endl;
            return 1;
        }
        std::cout << "sort success" << std::endl;
    }
    return 0;
}
