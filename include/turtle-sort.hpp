// turtle-sort.hpp -- drop-in C++ API of TurtleSort for AMD Instinct MI355X.
//
// Same names and call syntax as the reference header (turtle-sort.cuh):
//
//   Turtle::TurtleSort<Type, TrackIndex = true> sorter(maxElements, compress);
//   sorter.StartSorting(&values);            // or (&values, &indices)
//   double seconds = sorter.Sync();          // wall time incl. transfers
//   double s2 = sorter.MultiSort<Type, arrSize, blockSize>(numArrays, hostPtr);
//   bool c = sorter.MemoryCompressionSuccessful();   // always false here
//
// It is header-only, compiles with plain g++ (no HIP headers), and calls the
// C ABI of libturtle_hip.so (turtle_c.h); link with -lturtle_hip.
//
// Differences a caller can observe, all of them fixes:
//   - every key array is sorted and the index array follows the keys on
//     every size (the reference writes keys into the index on its merge path,
//     turtle-sort.cu:220, 255);
//   - equal keys keep their input order (the sort is stable), so the index
//     output is deterministic;
//   - MultiSort sorts the caller's data and every array, also a remainder
//     that is not a multiple of BlockSize (turtle-multi-sort.cuh:152, 200);
//   - sizes are checked: n > maxElements is an error, not an overrun
//     (turtle-sort.cuh:170);
//   - key types are the integer types of 1, 2, 4 and 8 bytes; `long` is 64
//     bits on Linux.
// Errors print a message and exit(status) like TurtleGlobals::gpuErrchk
// (turtle-globals.cuh:18-30); define TURTLE_THROW_ON_ERROR before including
// this header to get a std::runtime_error instead.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "turtle_c.h"

namespace TurtleGlobals {

inline void turtleAssert(int status, const char *what, const char *file, int line)
{
    if (status == TURTLE_OK) return;
    std::string msg = std::string("TurtleSort error: ") + turtle_status_string(status) + " (" + what + "): " +
                      turtle_last_error();
#ifdef TURTLE_THROW_ON_ERROR
    (void)file;
    (void)line;
    throw std::runtime_error(msg);
#else
    std::fprintf(stderr, "%s %s %d\n", msg.c_str(), file, line);
    std::exit(status);
#endif
}

#define TURTLE_CHECK(call) ::TurtleGlobals::turtleAssert((call), #call, __FILE__, __LINE__)

// RAII wall-clock timer with the semantics of the reference's
// TurtleGlobals::Bench (turtle-globals.cuh:33-58): nanoseconds into *target,
// or printed in seconds when target is null.
class Bench {
  public:
    explicit Bench(size_t *targetPtr) : target(targetPtr)
    {
        t1 = std::chrono::duration_cast<std::chrono::nanoseconds>(
            std::chrono::high_resolution_clock::now().time_since_epoch());
    }
    ~Bench()
    {
        auto t2 = std::chrono::duration_cast<std::chrono::nanoseconds>(
            std::chrono::high_resolution_clock::now().time_since_epoch());
        if (target)
            *target = (size_t)(t2.count() - t1.count());
        else
            std::cout << (t2.count() - t1.count()) / 1000000000.0 << " seconds" << std::endl;
    }

  private:
    size_t *target;
    std::chrono::nanoseconds t1;
};

}  // namespace TurtleGlobals

namespace Turtle {

// README.md:146 uses Turtle::Bench
using Bench = TurtleGlobals::Bench;

// key type code of T: the reference's ten instantiations (turtle-sort.cu:956-1174)
// plus 64-bit integers
template <typename T>
constexpr int dtypeOf()
{
    static_assert((std::is_integral<T>::value && !std::is_same<T, bool>::value) || std::is_same<T, float>::value ||
                      std::is_same<T, double>::value,
                  "TurtleSort keys are integer, float or double types");
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8, "unsupported key width");
    return std::is_same<T, float>::value    ? TURTLE_FLOAT32
           : std::is_same<T, double>::value ? TURTLE_FLOAT64
           : sizeof(T) == 1 ? (std::is_signed<T>::value ? TURTLE_INT8 : TURTLE_UINT8)
           : sizeof(T) == 2 ? (std::is_signed<T>::value ? TURTLE_INT16 : TURTLE_UINT16)
           : sizeof(T) == 4 ? (std::is_signed<T>::value ? TURTLE_INT32 : TURTLE_UINT32)
                            : (std::is_signed<T>::value ? TURTLE_INT64 : TURTLE_UINT64);
}

template <typename Type, bool TrackIndex = true>
struct TurtleSort {
  public:
    // turtle-sort.cuh:63-113
    TurtleSort(int maxElements, bool optInCompression = false) : maxN(maxElements)
    {
        TURTLE_CHECK(turtle_create(dtypeOf<Type>(), maxElements, TrackIndex ? 1 : 0, optInCompression ? 1 : 0, &h));
    }

    TurtleSort(const TurtleSort &) = delete;
    TurtleSort &operator=(const TurtleSort &) = delete;

    // turtle-sort.cuh:115-118
    bool MemoryCompressionSuccessful() { return turtle_compression_enabled(h) != 0; }

    // starts sorting on the GPU and returns; both vectors must stay alive and
    // unresized until Sync() (turtle-sort.cuh:120-184)
    void StartSorting(std::vector<Type> *arrayToSort, std::vector<int> *indicesToTrack = nullptr)
    {
        if (!TrackIndex) indicesToTrack = nullptr;
        if (indicesToTrack && indicesToTrack->size() != arrayToSort->size())
            TURTLE_CHECK(TURTLE_ERR_INVALID);
        toSort = arrayToSort;
        idTracker = indicesToTrack;
        TURTLE_CHECK(turtle_sort_start(h, arrayToSort->data(),
                                       indicesToTrack ? reinterpret_cast<int32_t *>(indicesToTrack->data()) : nullptr,
                                       (int64_t)arrayToSort->size()));
    }

    // waits, copies results back, returns seconds since StartSorting
    // (turtle-sort.cuh:186-258)
    double Sync()
    {
        double s = 0.0;
        TURTLE_CHECK(turtle_sort_sync(h, &s));
        toSort = nullptr;
        idTracker = nullptr;
        return s;
    }

    // sorts numArraysToSort consecutive arrays of ArrSize elements in place;
    // returns seconds including transfers (turtle-sort.cuh:260-276)
    template <typename T, int ArrSize, int BlockSize = 32, bool UseSharedMemory = false>
    double MultiSort(const int numArraysToSort, T *hostDataToSort)
    {
        static_assert(std::is_same<T, Type>::value, "MultiSort element type must be the sorter's Type");
        static_assert(ArrSize > 0 && BlockSize > 0, "ArrSize and BlockSize must be positive");
        double s = 0.0;
        TURTLE_CHECK(turtle_multisort(h, dtypeOf<Type>(), ArrSize, BlockSize, UseSharedMemory ? 1 : 0,
                                      numArraysToSort, hostDataToSort, &s));
        return s;
    }

    // MultiSort with index tracking, which the reference lists as
    // unsupported (README.md:46): index[i] moves with hostDataToSort[i],
    // stably; both arrays are sorted in place
    template <typename T, int ArrSize, int BlockSize = 32, bool UseSharedMemory = false>
    double MultiSort(const int numArraysToSort, T *hostDataToSort, int *index)
    {
        static_assert(std::is_same<T, Type>::value, "MultiSort element type must be the sorter's Type");
        static_assert(ArrSize > 0 && BlockSize > 0, "ArrSize and BlockSize must be positive");
        static_assert(TrackIndex, "MultiSort with an index needs a TurtleSort<Type, true>");
        double s = 0.0;
        TURTLE_CHECK(turtle_multisort_index(h, dtypeOf<Type>(), ArrSize, numArraysToSort, hostDataToSort,
                                            reinterpret_cast<int32_t *>(index), 0, &s));
        return s;
    }

    ~TurtleSort()
    {
        if (h) turtle_destroy(h);
    }

    // handle of the underlying C ABI object (for the device-resident calls)
    turtle_handle Handle() const { return h; }

  private:
    turtle_handle h = nullptr;
    int maxN;
    std::vector<Type> *toSort = nullptr;
    std::vector<int> *idTracker = nullptr;
};

}  // namespace Turtle
