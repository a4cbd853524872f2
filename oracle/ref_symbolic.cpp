// Checker harness (test infrastructure, never shipped): runs the reference's own
// symbolic analysis -- CSparse cs_amd / cs_etree / cs_post / cs_counts from
// cs.cpp and GetETree / GetSuperETree / GetRelaxedETree from etree.cpp, compiled
// from the reference sources by oracle/Makefile -- on one matrix and writes
// every intermediate array as raw little-endian int32 files.  The glue below
// restates the analyze steps of the reference driver (mdgpu.cu:1745-1875 and
// make_sym at mdgpu.cu:1630-1638) in order, so its outputs are the golden
// vectors the product's own symbolic phase must reproduce bit for bit.
//
// usage: ref_symbolic IN.bin OUTDIR AMD(0|1) MODE(plain|sn|rn) MAX_CEROS
// IN.bin: int32 n, int32 nnz, int32 Cp[n+1], int32 Ci[nnz]  (lower CSC incl.
// diagonal, rows sorted).  Values are irrelevant to the symbolic phase and
// are set to 1.

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cs.h"
#include "etree.h"

static void put(const std::string& dir, const char* name, const int* v, long n) {
    std::string path = dir + "/" + name + ".i32";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(2); }
    if (n > 0 && fwrite(v, sizeof(int), (size_t)n, f) != (size_t)n) { perror("fwrite"); exit(2); }
    fclose(f);
}

static int keep_offdiag(int i, int j, FLOTANTE, void*) { return i != j; }

// full symmetric pattern from the lower triangle (mdgpu.cu:1630-1638)
static cs* full_from_lower(cs* low) {
    cs* t = cs_transpose(low, 1);
    cs_fkeep(t, &keep_offdiag, NULL);
    cs* c = cs_add(low, t, 1, 1);
    cs_spfree(t);
    return c;
}

int main(int argc, char** argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s IN.bin OUTDIR AMD MODE MAX_CEROS\n", argv[0]);
        return 2;
    }
    const std::string out = argv[2];
    const int use_amd = atoi(argv[3]);
    const std::string mode = argv[4];
    const int max_ceros = atoi(argv[5]);

    FILE* f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    int n = 0, nnz = 0;
    if (fread(&n, 4, 1, f) != 1 || fread(&nnz, 4, 1, f) != 1) return 2;
    cs* low = cs_spalloc(n, n, nnz, 1, 0);
    if (fread(low->p, 4, (size_t)n + 1, f) != (size_t)n + 1) return 2;
    if (fread(low->i, 4, (size_t)nnz, f) != (size_t)nnz) return 2;
    fclose(f);
    for (int k = 0; k < nnz; k++) low->x[k] = 1;

    // fill-reducing ordering on the upper triangle, as the driver does
    std::vector<int> amd_perm(n);
    cs* a_low = NULL;
    if (use_amd) {
        cs* up = cs_transpose(low, 1);
        int* p = cs_amd(1, up);
        int* pinv = cs_pinv(p, n);
        cs* a_up = cs_symperm(up, pinv, 1);
        cs_spfree(up);
        a_low = cs_transpose(a_up, 1);
        cs_spfree(a_up);
        for (int k = 0; k < n; k++) amd_perm[k] = p[k];
        put(out, "amd_perm", p, n);
        free(p);
        free(pinv);
    } else {
        a_low = low;
        for (int k = 0; k < n; k++) amd_perm[k] = k;
    }
    cs* a = full_from_lower(a_low);

    int* parent = cs_etree(a, 0);
    int* post = cs_post(parent, n);
    int* count = cs_counts(a, parent, post, 0);
    put(out, "parent", parent, n);
    put(out, "post", post, n);
    put(out, "count", count, n);

    struct ETree** nodes = NULL;
    int nnodes = 0;
    GetETree(a, parent, post, count, &nodes, &nnodes);

    struct ETree** list = nodes;
    int nlist = nnodes;
    std::vector<int> perm_final(amd_perm);
    std::vector<int> count_final(count, count + n);
    if (mode == "sn") {
        GetSuperETree(nodes, nnodes, &list, &nlist);
    } else if (mode == "rn") {
        int* rperm = NULL;
        GetRelaxedETree(nodes, nnodes, max_ceros, count, &rperm, &list, &nlist);
        put(out, "relaxed_perm", rperm, n);
        int* rpinv = cs_pinv(rperm, n);
        cs* t1 = cs_transpose(a_low, 1);
        cs* t2 = cs_symperm(t1, rpinv, 1);
        cs_spfree(t1);
        cs* relaxed_low = cs_transpose(t2, 1);
        cs_spfree(t2);
        a_low = relaxed_low;
        for (int k = 0; k < n; k++) count_final[k] = count[rperm[k]];
        for (int k = 0; k < n; k++) perm_final[k] = amd_perm[rperm[k]];
        free(rpinv);
    } else if (mode != "plain") {
        fprintf(stderr, "bad mode %s\n", mode.c_str());
        return 2;
    }
    put(out, "perm_final", perm_final.data(), n);
    put(out, "count_final", count_final.data(), n);
    put(out, "alow_p", a_low->p, n + 1);
    put(out, "alow_i", a_low->i, a_low->p[n]);

    // node list in processing order: nodo, cols, rows, ceros, nHijos, child nodos
    std::vector<int> flat;
    flat.push_back(nlist);
    for (int k = 0; k < nlist; k++) {
        struct ETree* t = list[k];
        flat.push_back(t->nodo);
        flat.push_back(t->cols);
        flat.push_back(t->rows);
        flat.push_back(t->ceros);
        flat.push_back(t->nHijos);
        for (int h = 0; h < t->nHijos; h++) flat.push_back(t->hijos[h]->nodo);
    }
    put(out, "nodes", flat.data(), (long)flat.size());

    std::vector<int> lp(n + 1);
    long nz = 0;
    for (int k = 0; k < n; k++) { lp[k] = (int)nz; nz += count_final[k]; }
    lp[n] = (int)nz;
    put(out, "lp", lp.data(), n + 1);
    printf("n=%d nnz_low=%d nodes=%d lnz=%ld\n", n, nnz, nlist, nz);
    return 0;
}
