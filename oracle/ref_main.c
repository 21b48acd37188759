/*
 * ref_main.c -- process driver for the reference build in oracle/_ref.
 *
 * TEST INFRASTRUCTURE ONLY.  Stands in for SRC/mpi_rtm_iso2d.c on a single
 * rank: it calls the reference's rtm_routine(sim) for each simulation id in
 * [first, last] (SRC/mpi_rtm_iso2d.c:47-51) and prints the wall time of the
 * whole loop, as the reference does with MPI_Wtime (lines 46, 52, 54).
 *
 * usage: ref_main FIRST LAST      (run from a directory holding INPUTS/,
 *                                  OUTPUTS/ and TEMPORARY_FILES/)
 */
#include <stdio.h>
#include <stdlib.h>
#include <omp.h>

void rtm_routine(int simulationId);

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s FIRST LAST\n", argv[0]);
        return 2;
    }
    const int first = atoi(argv[1]), last = atoi(argv[2]);
    const double t0 = omp_get_wtime();
    for (int s = first; s <= last; s++) rtm_routine(s);
    const double t1 = omp_get_wtime();
    printf("REF_SECONDS %.6f THREADS %d\n", t1 - t0, omp_get_max_threads());
    return 0;
}
