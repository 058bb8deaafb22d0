// Matrix Market batch reader / writers of the drop-in interface.
#include "ReadWriteData.h"

#include <dirent.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "SolverResults.h"
#include "matrix.h"

namespace {

struct MMHeader {
    bool coordinate = false, symmetric = false;
    int rows = 0, cols = 0, entries = 0;
};

// banner "%%MatrixMarket matrix coordinate|array real|integer general|symmetric",
// comment lines, then the size line
MMHeader read_header(std::ifstream &f, const std::string &path)
{
    std::string line;
    if (!std::getline(f, line) || line.rfind("%%MatrixMarket", 0) != 0)
        throw std::runtime_error("not a Matrix Market file: " + path);
    std::string lower = line;
    std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
    MMHeader h;
    if (lower.find("matrix") == std::string::npos) throw std::runtime_error("not a matrix: " + path);
    h.coordinate = lower.find("coordinate") != std::string::npos;
    if (!h.coordinate && lower.find("array") == std::string::npos) throw std::runtime_error("unknown format: " + path);
    if (lower.find("complex") != std::string::npos || lower.find("pattern") != std::string::npos)
        throw std::runtime_error("only real / integer values are supported: " + path);
    h.symmetric = lower.find("symmetric") != std::string::npos;
    if (!h.symmetric && lower.find("general") == std::string::npos)
        throw std::runtime_error("only general / symmetric matrices are supported: " + path);
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream ss(line);
        if (h.coordinate) ss >> h.rows >> h.cols >> h.entries;
        else ss >> h.rows >> h.cols;
        if (!ss) throw std::runtime_error("bad size line: " + path);
        return h;
    }
    throw std::runtime_error("missing size line: " + path);
}

// CSR (rows sorted by column) of a coordinate file; symmetric files are expanded
void read_matrix(const std::string &path, int &n, std::vector<int> &rp, std::vector<int> &ci, std::vector<double> &v)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    const MMHeader h = read_header(f, path);
    if (!h.coordinate) throw std::runtime_error("matrix A must be in coordinate format: " + path);
    if (h.rows != h.cols) throw std::runtime_error("matrix is not square: " + path);
    n = h.rows;
    std::vector<std::tuple<int, int, double>> e;
    e.reserve(size_t(h.entries) * (h.symmetric ? 2 : 1));
    for (int k = 0; k < h.entries; k++) {
        int i, j;
        double x;
        if (!(f >> i >> j >> x)) throw std::runtime_error("short file: " + path);
        if (i < 1 || j < 1 || i > n || j > n) throw std::runtime_error("index out of range: " + path);
        e.emplace_back(i - 1, j - 1, x);
        if (h.symmetric && i != j) e.emplace_back(j - 1, i - 1, x);
    }
    std::sort(e.begin(), e.end(), [](const auto &a, const auto &b) {
        return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b) : std::get<1>(a) < std::get<1>(b);
    });
    for (size_t k = 1; k < e.size(); k++)
        if (std::get<0>(e[k]) == std::get<0>(e[k - 1]) && std::get<1>(e[k]) == std::get<1>(e[k - 1]))
            throw std::runtime_error("duplicate entry: " + path);
    rp.assign(n + 1, 0);
    ci.resize(e.size());
    v.resize(e.size());
    for (size_t k = 0; k < e.size(); k++) {
        rp[std::get<0>(e[k]) + 1]++;
        ci[k] = std::get<1>(e[k]);
        v[k] = std::get<2>(e[k]);
    }
    for (int i = 0; i < n; i++) rp[i + 1] += rp[i];
}

void read_vector(const std::string &path, int n, double *out)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    const MMHeader h = read_header(f, path);
    if (h.coordinate || h.cols != 1 || h.rows != n) throw std::runtime_error("b must be an n x 1 array: " + path);
    for (int i = 0; i < n; i++)
        if (!(f >> out[i])) throw std::runtime_error("short file: " + path);
}

}  // namespace

void FillSubdir(const std::string &main_dir, std::vector<std::string> &subdir)
{
    DIR *dr = opendir(main_dir.c_str());
    if (!dr) throw std::runtime_error("could not open directory " + main_dir);
    std::vector<std::string> names;
    while (dirent *de = readdir(dr)) {
        if (std::strcmp(de->d_name, ".") == 0 || std::strcmp(de->d_name, "..") == 0) continue;
        names.push_back(de->d_name);
    }
    closedir(dr);
    std::sort(names.begin(), names.end());
    for (const auto &s : names) subdir.push_back(main_dir + s);
}

void ReadData(const int Problem_Size, const std::vector<std::string> &subdir, PagedCSRMatrices &A_pages,
              PagedVectors &b_pages, const bool is_scaled)
{
    const int pages = Problem_Size * int(subdir.size());
    if (pages <= 0) throw std::runtime_error("no problems to read");
    const std::string fa = is_scaled ? "/A_scaled.mtx" : "/A.mtx", fb = is_scaled ? "/b_scaled.mtx" : "/b.mtx";
    int n = 0;
    std::vector<int> rp0, ci0;
    std::vector<std::vector<double>> vals(subdir.size());
    for (size_t d = 0; d < subdir.size(); d++) {
        int nd;
        std::vector<int> rp, ci;
        read_matrix(subdir[d] + fa, nd, rp, ci, vals[d]);
        if (d == 0) {
            n = nd; rp0 = rp; ci0 = ci;
        } else if (nd != n || rp != rp0 || ci != ci0) {
            throw std::runtime_error("sparsity pattern differs from the first problem: " + subdir[d] + fa);
        }
    }
    const int nnz = int(ci0.size());
    A_pages.DeallocateMemory(LOCATION::CPU);
    A_pages.SetNumPages(pages); A_pages.SetNumRows(n); A_pages.SetNumCols(n); A_pages.SetNumNz(nnz);
    A_pages.AllocateMemory(LOCATION::CPU);
    std::copy(rp0.begin(), rp0.end(), A_pages.GetPtrToCpuRowPtrs());
    std::copy(ci0.begin(), ci0.end(), A_pages.GetPtrToCpuColInd());
    b_pages.DeallocateMemory(LOCATION::CPU);
    b_pages.SetNumPages(pages); b_pages.SetNumElements(n);
    b_pages.AllocateMemory(LOCATION::CPU);
    std::vector<double> bd(n);
    for (int p = 0; p < pages; p++) {
        const size_t d = size_t(p) % subdir.size();
        std::copy(vals[d].begin(), vals[d].end(), A_pages.GetPtrToCpuValues() + size_t(p) * nnz);
        if (p < int(subdir.size())) read_vector(subdir[d] + fb, n, b_pages.GetPtrToCpuValues() + size_t(p) * n);
        else std::copy(b_pages.GetPtrToCpuValues() + d * n, b_pages.GetPtrToCpuValues() + (d + 1) * n,
                       b_pages.GetPtrToCpuValues() + size_t(p) * n);
    }
}

void Print_ans(const std::vector<std::string> &subdir, const PagedVectors &Ans_pages, const std::string &ans_name)
{
    if (!Ans_pages.ExistsCPU()) throw std::runtime_error("Print_ans: no host copy of the answers");
    const int M = Ans_pages.GetNumElements();
    for (size_t p = 0; p < subdir.size() && int(p) < Ans_pages.GetNumPages(); p++) {
        const std::string path = subdir[p] + "/" + ans_name;
        FILE *fp = std::fopen(path.c_str(), "w");
        if (!fp) throw std::runtime_error("cannot write " + path);
        std::fprintf(fp, "%%%%MatrixMarket matrix array real general\n%d %d\n", M, 1);
        for (int i = 0; i < M; i++) std::fprintf(fp, "%0.17lg\n", Ans_pages.GetPtrToCpuValues()[i + p * M]);
        std::fclose(fp);
    }
}

void PrintSolverResults(const std::vector<std::string> &subdir, const SolverResults &solver_results,
                        const std::string &f_name)
{
    if (!solver_results.ExistsCPU()) throw std::runtime_error("PrintSolverResults: no host copy of the results");
    FILE *fp = std::fopen(f_name.c_str(), "w");
    if (!fp) throw std::runtime_error("cannot write " + f_name);
    for (size_t i = 0; i < subdir.size() && int(i) < solver_results.GetNumPages(); i++)
        std::fprintf(fp, "%d   %f  \n", solver_results.GetPtrToCpuConvFlag()[i], solver_results.GetPtrToCpuIterCount()[i]);
    std::fclose(fp);
}
