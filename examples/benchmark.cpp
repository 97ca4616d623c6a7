// The README benchmark program (README.md:86-219) against the drop-in header:
// n keys + index via StartSorting/Sync, timed beside std::qsort, std::sort
// and std::sort(par_unseq) on {key, index} structs.  Keys are rand()-style
// draws in [0, 32767]; the index starts as iota (the README sets it to the
// key, which cannot detect a wrong index) and is checked to follow its key.
#include <algorithm>
#include <cstdlib>
#include <execution>
#include <iostream>
#include <vector>
#include "turtle-sort.hpp"

int main(int argc, char **argv)
{
    using Type = int;
    const int n = argc > 1 ? std::atoi(argv[1]) : 12 * 1024 * 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 3;
    Turtle::TurtleSort<Type> sort(n, false);
    std::vector<Type> hostData(n), original(n);
    std::vector<int> hostIndex(n);
    struct StdData { Type data; int index; };
    std::vector<StdData> backup(n), backup2(n), backup3(n);
    srand(777);
    for (int j = 0; j < reps; j++) {
        for (int i = 0; i < n; i++) {
            hostData[i] = rand() & 32767;
            original[i] = hostData[i];
            hostIndex[i] = i;
            backup[i] = backup2[i] = backup3[i] = StdData{hostData[i], i};
        }
        size_t t1, t2, t3, t4;
        {
            Turtle::Bench bench(&t1);
            sort.StartSorting(&hostData, &hostIndex);
            sort.Sync();
        }
        {
            Turtle::Bench bench(&t2);
            std::qsort(backup.data(), backup.size(), sizeof(StdData), [](const void *x, const void *y) {
                const int a = static_cast<const StdData *>(x)->data, b = static_cast<const StdData *>(y)->data;
                return a < b ? -1 : (a > b ? 1 : 0);
            });
        }
        {
            Turtle::Bench bench(&t3);
            std::sort(backup2.begin(), backup2.end(), [](auto &a, auto &b) { return a.data < b.data; });
        }
        {
            Turtle::Bench bench(&t4);
            std::sort(std::execution::par_unseq, backup3.begin(), backup3.end(),
                      [](auto &a, auto &b) { return a.data < b.data; });
        }
        std::cout << "gpu: " << t1 / 1e9 << "   std::qsort:" << t2 / 1e9 << "   std::sort:" << t3 / 1e9
                  << "   std::sort(par_unseq):" << t4 / 1e9 << std::endl;
        for (int i = 0; i < n; i++) {
            if (hostData[i] != backup2[i].data) {
                std::cout << "error: keys differ from std::sort at " << i << std::endl;
                return 1;
            }
            if (original[hostIndex[i]] != hostData[i] || (i > 0 && hostData[i] == hostData[i - 1] &&
                                                          hostIndex[i] <= hostIndex[i - 1])) {
                std::cout << "error: index calculation wrong at " << i << std::endl;
                return 1;
            }
        }
        std::cout << "quicksort (" << n << " elements) completed successfully" << std::endl;
    }
    return 0;
}
