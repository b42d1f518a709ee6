// main.cpp -- command-line front end with the reference's interface
// (main.cu:11-62):  PlexEnum <dataset> -k <k> -q <q> -t <thres> -tr <truss>
// It prints the reference's lines (main.cu:30,54; inc/graph.h:47;
// inc/host_funcs.h:115,871,873,1149,1156-1158), with a 64-bit plex count and one
// extra "hash:" line at the end.
#include "../../include/kplex.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

static void usage() { std::fprintf(stderr, "usage : ./PlexEnum <dataset> -k <k> -q <q> -t <thres> -tr <truss>\n"); }

static int check_inc(int i, int max)
{
    if (i == max) {
        usage();
        std::exit(1);
    }
    return i + 1;
}

int main(int argc, char *argv[])
{
    int i = 2;
    if (argc < 6) {
        usage();
        return 1;
    }
    std::printf("file: %s\n", argv[1]);
    kp_params p{};
    p.k = 0;
    p.q = 0;
    p.thres = 0.f;
    p.truss = 0;
    while (i < argc) {
        if (!std::strcmp(argv[i], "-k")) {
            i = check_inc(i, argc);
            p.k = std::atoi(argv[i]);
        } else if (!std::strcmp(argv[i], "-q")) {
            i = check_inc(i, argc);
            p.q = std::atoi(argv[i]);
        } else if (!std::strcmp(argv[i], "-t")) {
            i = check_inc(i, argc);
            p.thres = float(std::strtod(argv[i], nullptr));
        } else if (!std::strcmp(argv[i], "-tr")) {
            i = check_inc(i, argc);
            p.truss = std::atoi(argv[i]);
        } else {
            usage();
            return 1;
        }
        i++;
    }
    std::printf("k=%d,q=%d,thres=%f, truss: %d\n", p.k, p.q, p.thres, p.truss);
    kp_graph g{};
    int rc = kp_graph_load_bin(argv[1], &g);
    if (rc) {
        std::fprintf(stderr, "%s: %s\n", kp_strerror(rc), kp_last_error());
        return 1;
    }
    std::printf("n=%u, m=%llu (undirected)\n", g.n, (unsigned long long)g.m);
    kp_result r{};
    rc = kp_enumerate(&g, &p, &r);
    if (rc) {
        std::fprintf(stderr, "kp_enumerate: %s: %s\n", kp_strerror(rc), kp_last_error());
        kp_graph_free(&g);
        return 1;
    }
    // one "pn =" line per core peel: after the first core, and with the
    // truss on again after the second core (inc/host_funcs.h:115, :749, :823)
    std::printf("\npn = %u\n", r.pn_first);
    if (p.truss) std::printf("\npn = %u\n", r.pn);
    // the reference's lines around its host->device copy (inc/host_funcs.h:871,873)
    std::printf("Start copying graph to GPU....\n");
    std::printf("Done copying graph to GPU....\n");
    std::printf("Total Time Elapsed: %f ms\n", r.wall_ms);
    std::printf("Total Valid Blocks: %llu, Maximal k-Plexes: %llu\n", (unsigned long long)r.valid_roots,
                (unsigned long long)r.count);
    std::printf("Total tasks generated: %llu\n", (unsigned long long)r.tasks);
    std::printf("\nKernel Launch Successfully\n");
    std::printf("hash: %016llx%016llx\n", (unsigned long long)r.hash_b, (unsigned long long)r.hash_a);
    kp_result_free(&r);
    kp_graph_free(&g);
    return 0;
}
