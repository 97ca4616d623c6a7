// The README sample program (README.md:54-76) against the drop-in header.
// Expected output: 1 3 4 5 8 9
#include <iostream>
#include "turtle-sort.hpp"

int main()
{
    using Type = int;
    const int n = 1000;   // maximum array length to sort
    Turtle::TurtleSort<Type> sortVal(n);
    std::vector<Type> sample = {5, 4, 3, 9, 8, 1};
    sortVal.StartSorting(&sample);
    sortVal.Sync();
    for (auto &e : sample) std::cout << e << " ";
    std::cout << std::endl;
    std::cout << "Memory compression successful=" << sortVal.MemoryCompressionSuccessful() << std::endl;
    return 0;
}
