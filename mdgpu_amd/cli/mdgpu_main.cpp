// mdgpu command-line driver over the drop-in C++ interface (cs.h, etree.h,
// mdgpu.h): the flags and steps of the reference driver (mdgpu.cu:1666-1939)
// -- load a lower-triangular triplet file, optional AMD ordering, etree,
// postorder, counts, plain / fundamental / relaxed node tree, Factor() --
// followed by the solve and backward-error check the reference left after
// its exit(1) (mdgpu.cu:1941-1973), run here.
//
// usage: mdgpu [-gpu] [-amd] [-sn|-rn c] [-b bs] [-pad] [-symbolic] [-dump DIR] matrix
//   matrix     the reference's headerless "i j x" triplets (cs_load), or a
//              Matrix Market coordinate file (read by cs_load_mm)
//   -cpu       rejected: the numeric factorization runs on the GPU only
//   -symbolic  stop after the analysis (no GPU needed)
//   -dump DIR  write perm / parent / post / count / lp as raw int32 files
// The report prints the reference's line items (mdgpu.cu:1901-1930) from
// FactorGetTimes.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cs.h"
#include "../../include/etree.h"
#include "../../include/mdgpu.h"
#include "../../include/mdgpu_c.h"

static int offdiag(int i, int j, FLOTANTE, void*) { return i != j; }

// full symmetric matrix from its lower triangle (make_sym, mdgpu.cu:1630-1638)
static cs* make_sym(cs* A) {
    cs* AT = cs_transpose(A, 1);
    cs_fkeep(AT, &offdiag, NULL);
    cs* C = cs_add(A, AT, 1, 1);
    cs_spfree(AT);
    return C;
}

static FLOTANTE inf_norm(const FLOTANTE* x, int n) {
    FLOTANTE m = 0;
    for (int t = 0; t < n; t++) m = CS_MAX(m, fabs(x[t]));
    return m;
}

static void dump(const std::string& dir, const char* name, const int* v, int n) {
    std::string path = dir + "/" + name + ".i32";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return;
    fwrite(v, sizeof(int), n, f);
    fclose(f);
}

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
    bool amd = false, super = false, relaxed = false, symbolic_only = false;
    int max_ceros = 0;
    std::string dumpdir;
    if (argc < 2) {
        printf("mdgpu [-gpu] [-amd] [-sn|-rn c] [-b bs] [-pad] [-symbolic] [-dump DIR] matriz\n");
        return 1;
    }
    for (int t = 1; t < argc - 1; t++) {
        if (!strcmp(argv[t], "-gpu")) GPU = true;
        else if (!strcmp(argv[t], "-cpu")) {
            fprintf(stderr, "mdgpu: -cpu is not supported: this build factors on the GPU only "
                            "(the reference's CPU BLAS path, FactorAux_CPU, is not part of it)\n");
            return 1;
        }
        else if (!strcmp(argv[t], "-amd")) amd = true;
        else if (!strcmp(argv[t], "-sn")) super = true;
        else if (!strcmp(argv[t], "-rn")) { relaxed = true; max_ceros = atoi(argv[t + 1]); }
        else if (!strcmp(argv[t], "-b")) block_size = atoi(argv[t + 1]);
        else if (!strcmp(argv[t], "-pad")) pad = true;
        else if (!strcmp(argv[t], "-symbolic")) symbolic_only = true;
        else if (!strcmp(argv[t], "-dump")) dumpdir = argv[t + 1];
        else if (!strcmp(argv[t], "-help")) {
            printf("mdgpu [-gpu] [-amd] [-sn|-rn c] [-b bs] [-pad] [-symbolic] [-dump DIR] matriz\n");
            return 1;
        }
    }
    double t_load = now();
    FILE* f = fopen(argv[argc - 1], "r");
    if (!f) { perror(argv[argc - 1]); return 1; }
    char head[16] = {0};
    const bool mm = fread(head, 1, 14, f) == 14 && !strncmp(head, "%%MatrixMarket", 14);
    rewind(f);
    cs* T = mm ? cs_load_mm(f) : cs_load(f);
    fclose(f);
    if (!T || T->nz <= 0) { fprintf(stderr, "mdgpu: cannot read %s\n", argv[argc - 1]); return 1; }
    cs* C_low = cs_compress(T);
    cs_spfree(T);
    t_load = now() - t_load;
    const int n = C_low->n;

    double t_sym = now();
    cs* A_low;
    std::vector<int> amd_perm(n);
    for (int k = 0; k < n; k++) amd_perm[k] = k;
    if (amd) {
        cs* C_up = cs_transpose(C_low, 1);
        int* p = cs_amd(1, C_up);
        int* pinv = cs_pinv(p, n);
        cs* A_up = cs_symperm(C_up, pinv, 1);
        cs_spfree(C_up);
        A_low = cs_transpose(A_up, 1);
        cs_spfree(A_up);
        for (int k = 0; k < n; k++) amd_perm[k] = p[k];
        free(p);
        free(pinv);
    } else {
        A_low = cs_transpose(C_low, 1);       // copy
        cs* back = cs_transpose(A_low, 1);
        cs_spfree(A_low);
        A_low = back;
    }
    cs* A = make_sym(A_low);
    int* parent = cs_etree(A, 0);
    int* post = cs_post(parent, n);
    int* count = cs_counts(A, parent, post, 0);
    struct ETree** nodes = NULL;
    int nnodes = 0;
    if (!GetETree(A, parent, post, count, &nodes, &nnodes)) return 1;
    printf("\nCant Nodos = %i\n", nnodes);
    struct ETree** list = nodes;
    int nlist = nnodes;
    if (super) {
        GetSuperETree(nodes, nnodes, &list, &nlist);
        printf("\nCant SuperNodos = %i\n", nlist);
    }
    std::vector<int> perm(amd_perm);
    if (relaxed) {
        int* rperm = NULL;
        GetRelaxedETree(nodes, nnodes, max_ceros, count, &rperm, &list, &nlist);
        if (!list) return 1;
        printf("\nCant RelaxedNodos = %i\n", nlist);
        int* rpinv = cs_pinv(rperm, n);
        cs* t1 = cs_transpose(A_low, 1);
        cs* t2 = cs_symperm(t1, rpinv, 1);
        cs_spfree(t1);
        cs_spfree(A_low);
        A_low = cs_transpose(t2, 1);
        cs_spfree(t2);
        std::vector<int> c2(n);
        cs_pvec(rperm, count, c2.data(), n);
        memcpy(count, c2.data(), sizeof(int) * n);
        for (int k = 0; k < n; k++) perm[k] = amd_perm[rperm[k]];
        free(rperm);
        free(rpinv);
    }
    cs* spL = cs_spalloc(n, n, 1, 0, 0);
    free(spL->p);
    spL->p = (int*)malloc(sizeof(int) * (n + 1));
    int nz = 0;
    for (int k = 0; k < n; k++) { spL->p[k] = nz; nz += count[k]; }
    spL->p[n] = nz;
    spL->nzmax = nz;
    free(spL->i);
    spL->i = (int*)calloc(nz > 0 ? nz : 1, sizeof(int));
    spL->x = NULL;
    printf("L.nz = %i\n", nz);
    t_sym = now() - t_sym;
    if (!dumpdir.empty()) {
        dump(dumpdir, "amd_perm", amd_perm.data(), n);
        dump(dumpdir, "perm_final", perm.data(), n);
        dump(dumpdir, "parent", parent, n);
        dump(dumpdir, "post", post, n);
        dump(dumpdir, "count_final", count, n);
        dump(dumpdir, "lp", spL->p, n + 1);
        dump(dumpdir, "alow_p", A_low->p, n + 1);
        dump(dumpdir, "alow_i", A_low->i, A_low->p[n]);
    }
    printf("Load           = %.6f\n", t_load);
    printf("Symbolic       = %.6f\n", t_sym);
    if (symbolic_only) return 0;

    Factor(A_low, spL, list, nlist);
    if (!spL->x) { fprintf(stderr, "Factor failed: %s\n", mdgpu_last_error()); return 2; }
    FactorTimes ft;
    FactorGetTimes(&ft);
    // the reference's report (mdgpu.cu:1901-1930); its CPU BLAS items are 0
    printf("\nMFlops     = %.2f\n", ft.mflops);
    printf("MFlops/s   = %.3f\n\n", ft.factor_aux > 0 ? ft.mflops / ft.factor_aux : 0.0);
    printf("Factor         = %.6f Factorizacion numerica total\n", ft.factor);
    printf("FactorAux      = %.6f Factorizacion numerica solo calculo\n", ft.factor_aux);
    printf("  BLAS CPU     = %.6f\n", 0.0);
    printf("  BLAS GPU     = %.6f\n", ft.blas_gpu);
    printf("    TRSM GPU   = %.6f\n", ft.trsm_gpu);
    printf("    GEMM GPU   = %.6f\n", ft.gemm_gpu);
    printf("  Otros        = %.6f\n", ft.other_gpu);
    printf("Memcpy         = %.6f\n", ft.memcpy_x + ft.memcpy_l);
    printf("  Memcpy2      = %.6f Copia de L GPU->CPU\n", ft.memcpy_l);
    printf("  MemcpyX      = %.6f Copia inicial de A CPU->GPU\n", ft.memcpy_x);
    printf("Merge          = %.6f\n", ft.merge);
    printf("GetFk          = %.6f Formar el frente\n", ft.getfk);
    printf("ExtendAdd      = %.6f Suma extendida\n", ft.extend_add);
    printf("Malloc GPU     = %.6f\n", ft.malloc_gpu);
    printf("Free GPU       = %.6f\n\n", ft.free_gpu);

    // solve + backward error (mdgpu.cu:1941-1973), b_i = 1 + i/n
    std::vector<FLOTANTE> b(n), x(n), x2(n), resid(n);
    for (int t = 0; t < n; t++) b[t] = 1 + ((FLOTANTE)t) / n;
    cs_pvec(perm.data(), b.data(), x.data(), n);
    cs_lsolve(spL, x.data());
    cs_ltsolve(spL, x.data());
    cs_ipvec(perm.data(), x.data(), x2.data(), n);
    for (int t = 0; t < n; t++) resid[t] = -b[t];
    cs* Cfull = make_sym(C_low);
    cs_gaxpy(Cfull, x2.data(), resid.data());
    const double eta = inf_norm(resid.data(), n) / (cs_norm(Cfull) * inf_norm(x2.data(), n) + inf_norm(b.data(), n));
    printf("resid: %8.2e\n", eta);
    // FLOTANTE = float factors and solves in float: its backward error is
    // of the order of float's rounding
    return eta < (sizeof(FLOTANTE) == sizeof(float) ? 1e-5 : 1e-10) ? 0 : 3;
}
