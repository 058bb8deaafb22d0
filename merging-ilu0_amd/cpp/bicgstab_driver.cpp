// Driver of the drop-in C++ interface, in the manner of the reference's
// main_bicgstab_*.cpp programs: loads a batch from a raw binary file or a
// Matrix Market tree, calls one of the reference-named entry points and
// writes x and the per-page results.
//
//   bicgstab_driver <merged|segregated|ilu|parilu|ilu_isai|parilu_isai|general_isai> in.bin out.bin
//   bicgstab_driver <ilu0|parilu0|isai_general|isai_lower|isai_upper> in.bin out.bin
//   bicgstab_driver <variant> --mtx <category_dir/> <problem_size> <1 for scaled> [out.bin]
//   bicgstab_driver convert --mtx <category_dir/> <problem_size> <1 for scaled> out.bin
//
// --mtx reads the batch from a tree of Matrix Market problems (FillSubdir /
// ReadData: Problem_Size copies of every subdirectory's A.mtx / b.mtx); the
// solvers then write x_gpu_*.mtx into each problem directory and
// Results_<variant>_bicgstab_<category>.txt (convergence flag, iterations per
// problem) into the working directory.  convert writes the batch read from
// the tree in the in.bin layout and touches no GPU.
//
// in.bin:  int32 n, nnz, pages; int32 row_ptrs[n+1], col_ind[nnz];
//          float64 values[pages*nnz], b[pages*n]
// out.bin (solvers): float64 x[pages*n]; float32 iter[pages]; int32 conv[pages];
//          float64 iter_res[pages], true_res[pages]; float64 PST, IST, OET
// out.bin (ilu0 / ilu0_app1..4 / parilu0, ILU_0_Factorization_Gpu with that
//          approach / ParILU_0_Factorization_Gpu with 5 sweeps): int32 l_nnz,
//          u_nnz; int32 L row_ptrs[n+1], col_ind;
//          U row_ptrs, col_ind; float64 L values[pages*l_nnz], U values
// out.bin (isai_*, GenerateISAI_gpu power 1 of the input's typed view):
//          float64 values[pages*nnz]
// out.bin (isai2_general|isai2_lower|isai2_upper, power 2): int32 nnz_M;
//          int32 row_ptrs[n+1], col_ind[nnz_M]; float64 values[pages*nnz_M]
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "ILU_0.h"
#include "PrecondBiCGSTAB.h"
#include "ReadWriteData.h"
#include "SolverResults.h"
#include "isai.h"
#include "matrix.h"
#include "parILU_0.h"

template <typename T>
static bool rd(FILE *f, T *p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }
template <typename T>
static bool wr(FILE *f, const T *p, size_t n) { return std::fwrite(p, sizeof(T), n, f) == n; }

static void load_bin(const char *path, PagedCSRMatrices &A, PagedVectors &b)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    int hdr[3];
    bool ok = rd(f, hdr, 3);
    const int n = hdr[0], nnz = hdr[1], pages = hdr[2];
    if (ok) {
        A.SetNumPages(pages); A.SetNumRows(n); A.SetNumCols(n); A.SetNumNz(nnz);
        A.AllocateMemory(LOCATION::CPU);
        b.SetNumPages(pages); b.SetNumElements(n);
        b.AllocateMemory(LOCATION::CPU);
        ok = rd(f, A.GetPtrToCpuRowPtrs(), n + 1) && rd(f, A.GetPtrToCpuColInd(), nnz) &&
             rd(f, A.GetPtrToCpuValues(), (size_t)pages * nnz) && rd(f, b.GetPtrToCpuValues(), (size_t)pages * n);
    }
    std::fclose(f);
    if (!ok) throw std::runtime_error(std::string("short input ") + path);
}

static bool write_bin(const char *path, const PagedCSRMatrices &A, const PagedVectors &b)
{
    FILE *o = std::fopen(path, "wb");
    if (!o) return false;
    const int n = A.GetNumRows(), nnz = A.GetNumNz(), pages = A.GetNumPages();
    const int hdr[3] = {n, nnz, pages};
    const bool ok = wr(o, hdr, 3) && wr(o, A.GetPtrToCpuRowPtrs(), n + 1) && wr(o, A.GetPtrToCpuColInd(), nnz) &&
                    wr(o, A.GetPtrToCpuValues(), (size_t)pages * nnz) && wr(o, b.GetPtrToCpuValues(), (size_t)pages * n);
    std::fclose(o);
    return ok;
}

int main(int argc, char **argv)
{
    const bool mtx = argc >= 6 && std::strcmp(argv[2], "--mtx") == 0;
    if (!(argc == 4 || (mtx && argc <= 7))) {
        std::fprintf(stderr, "usage: %s <variant> in.bin out.bin\n"
                             "       %s <variant> --mtx <category_dir/> <problem_size> <1 for scaled> [out.bin]\n",
                     argv[0], argv[0]);
        return 2;
    }
    const std::string variant = argv[1];
    const char *out_path = mtx ? (argc == 7 ? argv[6] : nullptr) : argv[3];
    PagedCSRMatrices A;
    PagedVectors b;
    std::vector<std::string> subdir;
    std::string category;
    try {
        if (mtx) {
            std::string dir = argv[3];
            if (dir.empty() || dir.back() != '/') dir += '/';
            const size_t cut = dir.find_last_of('/', dir.size() - 2);
            category = dir.substr(cut == std::string::npos ? 0 : cut + 1);
            category.pop_back();
            FillSubdir(dir, subdir);
            ReadData(std::stoi(argv[4]), subdir, A, b, std::stoi(argv[5]) == 1);
        } else {
            load_bin(argv[2], A, b);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "reading the batch failed: %s\n", e.what());
        return 2;
    }
    const int n = A.GetNumRows(), nnz = A.GetNumNz(), pages = A.GetNumPages();
    const bool is_scaled = mtx && std::stoi(argv[5]) == 1;
    if (variant == "convert") {
        if (!out_path || !write_bin(out_path, A, b)) { std::fprintf(stderr, "convert: cannot write\n"); return 2; }
        std::printf("convert: %zu problems x %s = %d pages n=%d nnz=%d\n", subdir.size(), mtx ? argv[4] : "1", pages,
                    n, nnz);
        return 0;
    }
    A.AllocateMemory(LOCATION::GPU);
    A.CopyFromCpuToGpu();
    b.AllocateMemory(LOCATION::GPU);
    b.CopyFromCpuToGpu();
    // ilu0_app<k>: ILU_0_Factorization_Gpu with approach k (1..4; "ilu0" = 1)
    const bool ilu_app = variant.size() == 9 && variant.rfind("ilu0_app", 0) == 0 && variant[8] >= '1' &&
                         variant[8] <= '4';
    if (variant == "ilu0" || ilu_app || variant == "parilu0" || variant.rfind("isai_", 0) == 0 ||
        variant.rfind("isai2_", 0) == 0) {
        FILE *o = out_path ? std::fopen(out_path, "wb") : nullptr;
        if (!o) { std::perror("out.bin"); return 2; }
        bool ok = true;
        try {
            if (variant.rfind("isai", 0) == 0) {
                PagedCSRMatrices M;
                const bool p2 = variant.rfind("isai2_", 0) == 0;
                const std::string ty = variant.substr(p2 ? 6 : 5);
                const mat_type t = ty == "lower" ? mat_type::lower_tri
                                 : ty == "upper" ? mat_type::upper_tri : mat_type::general;
                GenerateISAI_gpu(M, A, t, p2 ? 2 : 1);
                M.CopyFromGpuToCpu();
                const int mnnz = M.GetNumNz();
                if (p2)
                    ok = wr(o, &mnnz, 1) && wr(o, M.GetPtrToCpuRowPtrs(), n + 1) && wr(o, M.GetPtrToCpuColInd(), mnnz);
                ok = ok && wr(o, M.GetPtrToCpuValues(), (size_t)pages * mnnz);
            } else {
                PagedCSRMatrices L, U;
                if (variant == "ilu0") ILU_0_Factorization_Gpu(A, L, U, 1);
                else if (ilu_app) ILU_0_Factorization_Gpu(A, L, U, variant[8] - '0');
                else ParILU_0_Factorization_Gpu(A, L, U, 5);
                L.CopyFromGpuToCpu();
                U.CopyFromGpuToCpu();
                const int sz[2] = {L.GetNumNz(), U.GetNumNz()};
                ok = wr(o, sz, 2) && wr(o, L.GetPtrToCpuRowPtrs(), n + 1) && wr(o, L.GetPtrToCpuColInd(), sz[0]) &&
                     wr(o, U.GetPtrToCpuRowPtrs(), n + 1) && wr(o, U.GetPtrToCpuColInd(), sz[1]) &&
                     wr(o, L.GetPtrToCpuValues(), (size_t)pages * sz[0]) &&
                     wr(o, U.GetPtrToCpuValues(), (size_t)pages * sz[1]);
            }
        } catch (const std::exception &e) {
            std::fprintf(stderr, "%s failed: %s\n", variant.c_str(), e.what());
            std::fclose(o);
            return 1;
        }
        std::fclose(o);
        std::printf("%s: %d pages n=%d nnz=%d\n", variant.c_str(), pages, n, nnz);
        return ok ? 0 : 1;
    }
    PagedVectors x;
    x.SetNumPages(pages); x.SetNumElements(n);
    x.AllocateMemory(LOCATION::CPU);
    SolverResults res(pages, ATOL, CPU_EXISTENCE::EXISTENT, GPU_EXISTENCE::EXISTENT);
    double PST = 0, IST = 0, OET = 0;
    try {
        if (variant == "merged")
            Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_merged_Gpu(subdir, A, b, x, is_scaled, res);
        else if (variant == "segregated")
            Batched_conv_ILU_app1_Preconditioned_BiCGSTAB_segregated_Gpu(subdir, A, b, x, is_scaled, res);
        else if (variant == "ilu" || variant == "parilu")
            Batched_ILU_Preconditioned_BiCGSTAB_Gpu(subdir, A, b, x, is_scaled, res, variant == "parilu", 1, 5, PST, IST, OET);
        else if (variant == "ilu_isai" || variant == "parilu_isai")
            Batched_ILU_ISAI_Preconditioned_BiCGSTAB_Gpu(subdir, A, b, x, is_scaled, res, variant == "parilu_isai", 5, PST,
                                                         IST, OET);
        else if (variant == "general_isai")
            Batched_General_ISAI_Preconditioned_BiCGSTAB_Gpu(subdir, A, b, x, is_scaled, res, PST, IST, OET);
        else {
            std::fprintf(stderr, "unknown variant %s\n", variant.c_str());
            return 2;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "solve failed: %s\n", e.what());
        return 1;
    }
    if (OET == 0) OET = IST = res.GetTimeTaken();
    try {
        if (mtx) PrintSolverResults(subdir, res, "Results_" + variant + "_bicgstab_" + category + ".txt");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "writing the results failed: %s\n", e.what());
        return 1;
    }
    bool ok = true;
    const double t[3] = {PST, IST, OET};
    FILE *o = out_path ? std::fopen(out_path, "wb") : nullptr;
    if (out_path && !o) { std::perror("out.bin"); return 2; }
    if (o) ok = wr(o, x.GetPtrToCpuValues(), (size_t)pages * n) && wr(o, res.GetPtrToCpuIterCount(), pages) &&
                    wr(o, res.GetPtrToCpuConvFlag(), pages) && wr(o, res.GetPtrToCpuIterResNorm(), pages) &&
                    wr(o, res.GetPtrToCpuTrueResNorm(), pages) && wr(o, t, 3);
    if (o) std::fclose(o);
    std::printf("%s: %d pages n=%d nnz=%d, PST %.3f ms IST %.3f ms OET %.3f ms\n", variant.c_str(), pages, n, nnz,
                PST, IST, OET);
    return ok ? 0 : 1;
}
