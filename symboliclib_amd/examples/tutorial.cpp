// The tutorial chain of SymbolicLib (examples/Tutorial/main.cpp:17-68 of the reference), written
// against this library's C++ surface: a scalar expression evaluated and differentiated, then
// B = A^T A + A built from makeSymbolic(A) and evaluated by ComputeUnit on the CPU device
// (VecWidth / NumThreads) and on the GPU (UseHIP, ThreadsPerBlock).
//
//   tutorial <matrix.mtx> [results.bin]
//
// Prints the scalar values, the difference of each unit's result to A^T A + A computed directly
// in double, and writes both units' results to results.bin when given:
//   int32 n, int32 nnz, int32 outer[n + 1], int32 inner[nnz], float64 cpu[nnz], float64 hip[nnz].
// Exits non-zero when the patterns differ or a unit fails (the GPU unit needs a GPU).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ComputeUnit.hpp"
#include "ContainerSupport.hpp"
#include "Sparse.hpp"
#include "Symbolic.hpp"
#include "SymbolicDifferentiation.hpp"

using namespace Sym;

namespace {

// MatrixMarket "coordinate real general", 1-based; duplicate entries are summed (the behaviour
// of Eigen::loadMarket, which goes through setFromTriplets).
SparseMatrix<double> loadMarket(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    std::getline(in, line);
    if (line.find("coordinate") == std::string::npos) throw std::runtime_error(path + ": not a coordinate matrix");
    while (std::getline(in, line) && !line.empty() && line[0] == '%') {
    }
    int rows = 0, cols = 0;
    long long nnz = 0;
    std::istringstream(line) >> rows >> cols >> nnz;
    std::vector<Triplet<double>> entries;
    entries.reserve((size_t)nnz);
    for (long long k = 0; k < nnz; ++k) {
        int r, c;
        double v;
        if (!(in >> r >> c >> v)) throw std::runtime_error(path + ": truncated");
        entries.emplace_back(r - 1, c - 1, v);
    }
    SparseMatrix<double> A(rows, cols);
    A.setFromTriplets(entries.begin(), entries.end());
    return A;
}

bool samePattern(const SparseMatrix<double>& a, const SparseMatrix<double>& b) {
    return a.rows() == b.rows() && a.cols() == b.cols() && a.outer() == b.outer() && a.inner() == b.inner();
}

double difference(const SparseMatrix<double>& a, const SparseMatrix<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.nonZeros(); ++i) {
        const double d = a.valuePtr()[i] - b.valuePtr()[i];
        s += d * d;
    }
    return std::sqrt(s);
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cerr << "usage: tutorial <matrix.mtx> [results.bin]\n";
        return 2;
    }
    try {
        // variables {id, group}; an expression of them, its value and its gradient at (1, 2)
        Symbolic a(0, 0);
        Symbolic b(1, 0);
        Symbolic c = a + b * sqrt(a * b);
        std::printf("c(1, 2) = %.6g\n", evaluate(c, std::vector<double>{1., 2.}));  // 3.82843
        auto dc = differentiate(c, std::vector<Symbolic>{a, b});
        std::printf("dc/da(1, 2) = %.6g\n", evaluate(dc[0], std::vector<double>{1., 2.}));  // 2.41421
        std::printf("dc/db(1, 2) = %.6g\n", evaluate(dc[1], std::vector<double>{1., 2.}));  // 2.12132

        // B = A^T A + A with A's values as the program's input
        SparseMatrix<double> A = loadMarket(argv[1]);
        SparseMatrix<Symbolic> AS = makeSymbolic(A, 0);
        SparseMatrix<Symbolic> BS = AS.transpose() * AS + AS;

        const SparseMatrix<double> B = A.transpose() * A + A;
        SparseMatrix<double> Bcpu = B, Bhip = B;
        for (size_t i = 0; i < B.nonZeros(); ++i) Bcpu.valuePtr()[i] = Bhip.valuePtr()[i] = 0.0;
        if (BS.nonZeros() != B.nonZeros()) throw std::runtime_error("symbolic and numeric patterns differ");

        // the CPU device: four doubles per vector, eight threads
        ComputeUnit<double> unit(Device(VecWidth(4), NumThreads(8)), AS, BS);
        unit.compile().execute(A);
        unit.getResults(Bcpu);
        unit.close();
        std::printf("difference: %.6g\n", difference(B, Bcpu));

        // the GPU: the same program for the HIP device
        ComputeUnit<double> unitHIP(Device(UseHIP(), ThreadsPerBlock(256)), AS, BS);
        unitHIP.compile().execute(A).getResults(Bhip);
        std::printf("difference hip: %.6g\n", difference(B, Bhip));
        if (!samePattern(B, Bcpu) || !samePattern(B, Bhip)) throw std::runtime_error("result patterns differ");

        if (argc > 2) {
            std::FILE* f = std::fopen(argv[2], "wb");
            if (!f) throw std::runtime_error(std::string("cannot write ") + argv[2]);
            const int n = B.cols(), nnz = (int)B.nonZeros();
            bool ok = std::fwrite(&n, sizeof n, 1, f) == 1 && std::fwrite(&nnz, sizeof nnz, 1, f) == 1;
            ok = ok && std::fwrite(B.outer().data(), sizeof(int), (size_t)n + 1, f) == (size_t)n + 1;
            ok = ok && std::fwrite(B.inner().data(), sizeof(int), (size_t)nnz, f) == (size_t)nnz;
            ok = ok && std::fwrite(Bcpu.valuePtr(), sizeof(double), (size_t)nnz, f) == (size_t)nnz;
            ok = ok && std::fwrite(Bhip.valuePtr(), sizeof(double), (size_t)nnz, f) == (size_t)nnz;
            ok = std::fclose(f) == 0 && ok;
            if (!ok) throw std::runtime_error(std::string("writing ") + argv[2] + " failed");
        }
    } catch (const std::exception& e) {
        std::cerr << "tutorial: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
