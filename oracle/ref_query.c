/* ref_query -- times one SQL query on the reference engine through its C API
 * (include/slothdb/api/slothdb.h:46-104): an in-memory database, the query run
 * `reps` times, each run's wall time from slothdb_query's call to its return
 * (plan, execution and the result materialized in the engine; nothing is
 * printed). Benchmark infrastructure: linked against the reference's own
 * sources by oracle/Makefile (target cli), never part of the product.
 *
 *   ref_query <reps> <sql>
 *
 * Prints one line per run: "run <i> <seconds> <rows>", then "ok". Exit status
 * 1 with the engine's message on error. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "slothdb/api/slothdb.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: ref_query <reps> <sql>\n");
        return 2;
    }
    const int reps = atoi(argv[1]);
    slothdb_database *db = NULL;
    slothdb_connection *conn = NULL;
    if (slothdb_open(NULL, &db) != SLOTHDB_OK || slothdb_connect(db, &conn) != SLOTHDB_OK) {
        fprintf(stderr, "ref_query: cannot open an in-memory database\n");
        return 1;
    }
    for (int i = 0; i < reps; i++) {
        slothdb_result *res = NULL;
        const double t0 = now_s();
        const slothdb_status st = slothdb_query(conn, argv[2], &res);
        const double dt = now_s() - t0;
        if (st != SLOTHDB_OK) {
            fprintf(stderr, "ref_query: %s\n", res ? slothdb_result_error(res) : "query failed");
            if (res) slothdb_free_result(res);
            return 1;
        }
        printf("run %d %.6f %llu\n", i, dt, (unsigned long long)slothdb_row_count(res));
        fflush(stdout);
        slothdb_free_result(res);
    }
    slothdb_disconnect(conn);
    slothdb_close(db);
    printf("ok\n");
    return 0;
}
