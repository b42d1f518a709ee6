/* kpo_cli -- command-line driver for the CPU oracle (test infrastructure).
 * usage: kpo_cli <graph.bin> -k K -q Q [-tr 0|1] [-j threads] [-list] [-range b e] */
#include "kp_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
    if (argc < 2) { fprintf(stderr, "usage: kpo_cli <graph.bin> -k K -q Q [-tr 0|1] [-j T] [-list] [-range b e]\n"); return 1; }
    int k = 2, q = 4, tr = 0, j = 0, list = 0;
    long rb = 0, re = -1;
    for (int i = 2; i < argc; i++) {
        if (!strcmp(argv[i], "-k") && i + 1 < argc) k = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-q") && i + 1 < argc) q = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-tr") && i + 1 < argc) tr = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-j") && i + 1 < argc) j = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-list")) list = 1;
        else if (!strcmp(argv[i], "-range") && i + 2 < argc) { rb = atol(argv[++i]); re = atol(argv[++i]); }
        else { fprintf(stderr, "bad arg %s\n", argv[i]); return 1; }
    }
    kpo_graph g;
    int rc = kpo_load_bin(argv[1], &g);
    if (rc) { fprintf(stderr, "load failed %d\n", rc); return 2; }
    kpo_result r;
    rc = kpo_enumerate_range(&g, k, q, tr, j, list ? KPO_LIST : 0, rb, re, &r);
    if (rc) { fprintf(stderr, "enumerate failed %d\n", rc); return 3; }
    printf("{\"n\": %u, \"m\": %llu, \"pn\": %u, \"pm\": %llu, \"count\": %llu, \"tasks\": %llu, \"valid_roots\": %llu, "
           "\"hash_a\": \"%016llx\", \"hash_b\": \"%016llx\", \"alg_bytes\": %llu, \"adj_bytes\": %llu, "
           "\"max_n\": %u, \"max_l\": %u, \"wall_ms\": %.3f, \"preprocess_ms\": %.3f}\n",
           g.n, (unsigned long long)g.m, r.pn, (unsigned long long)r.pm, (unsigned long long)r.count,
           (unsigned long long)r.tasks, (unsigned long long)r.valid_roots,
           (unsigned long long)r.hash_a, (unsigned long long)r.hash_b, (unsigned long long)r.alg_bytes,
           (unsigned long long)r.adj_bytes, r.max_n, r.max_l, r.wall_ms, r.preprocess_ms);
    if (list) {
        for (uint64_t p = 0; p < r.nlisted; p++) {
            for (uint64_t t = r.list_off[p]; t < r.list_off[p + 1]; t++)
                printf("%s%u", t == r.list_off[p] ? "" : " ", r.list_ids[t]);
            printf("\n");
        }
    }
    kpo_result_free(&r);
    kpo_graph_free(&g);
    return 0;
}
