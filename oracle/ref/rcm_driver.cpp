// Driver around the reference RCM baseline (rcm/src/rcm.cpp, rcm/src/queue.cpp),
// built by oracle/Makefile into oracle/_ref/rcm_driver.  Reads "n nz" then nz
// lines "row col" on stdin and prints the R_Cuthill_Mckee permutation.
#include <cstdio>
#include <vector>
#include "rcm.h"

int main()
{
    int n = 0, nz = 0;
    if (std::scanf("%d %d", &n, &nz) != 2) return 1;
    std::vector<int> rows(nz), cols(nz);
    for (int i = 0; i < nz; ++i)
        if (std::scanf("%d %d", &rows[i], &cols[i]) != 2) return 2;
    std::vector<unsigned int> perm(n);
    R_Cuthill_Mckee(n, nz, rows.data(), cols.data(), perm.data());
    for (int i = 0; i < n; ++i) std::printf("%u\n", perm[i]);
    return 0;
}
