// Front-ends that build the symbolic inputs/outputs of the three benchmark workloads, written
// against the drop-in API exactly as a user of the reference would (Eigen replaced by the
// build's own containers):
//   * A^T A + A            examples/Tutorial/main.cpp:36-43
//   * cotangent Laplacian  examples/BuildCotan/main.cpp:39-76
//   * Baraff-Witkin cloth  examples/BaraffWitkinAD/stretch_shear_forces_impl.hpp:9-73 (energy,
//     rest-shape precomputation), main.cpp:90-98 (AD Hessian), plus the constant quadratic
//     bending stiffness of examples/BaraffWitkin/implementation/bend_quadratic_forces.cpp:54-155.
// The scalar type NT is double (direct evaluation) or Sym::Symbolic (recorded expressions).
#pragma once

#include <array>
#include <cmath>
#include <map>
#include <stdexcept>
#include <vector>

#include "../sym/ContainerSupport.hpp"
#include "../sym/Sparse.hpp"
#include "../sym/Symbolic.hpp"
#include "../sym/SymbolicDifferentiation.hpp"

namespace Sym {
namespace workloads {

template <class T>
struct Vec3 {
    T x, y, z;
    Vec3 operator-(const Vec3& b) const { return {x - b.x, y - b.y, z - b.z}; }
    Vec3 operator+(const Vec3& b) const { return {x + b.x, y + b.y, z + b.z}; }
    T dot(const Vec3& b) const { return x * b.x + y * b.y + z * b.z; }
    Vec3 cross(const Vec3& b) const { return {y * b.z - z * b.y, z * b.x - x * b.z, x * b.y - y * b.x}; }
    T squaredNorm() const { return x * x + y * y + z * z; }
};

template <class T>
T norm(const Vec3<T>& v) {
    using std::sqrt;
    return sqrt(v.squaredNorm());
}

template <class T>
Vec3<T> scaled(const Vec3<T>& v, const T& s) {
    return {v.x * s, v.y * s, v.z * s};
}

// Row v of an n x 3 column-major matrix.
template <class T>
Vec3<T> row(const DenseMatrix<T>& X, int v) {
    return {X(v, 0), X(v, 1), X(v, 2)};
}

using Faces = std::vector<std::array<int, 3>>;

/* A^T A + A ******************************************************************/

template <class T>
SparseMatrix<T> ataPlusA(const SparseMatrix<T>& A) {
    return A.transpose() * A + A;
}

/* cotangent Laplacian ********************************************************/

// L(j1, j2) += w, L(j1, j1) -= w for the cotangent weight w = (e0 . e1) / |e0 x e1| of the
// corner opposite edge (j1, j2), four triplets per corner.
template <class T>
void buildCotan(const DenseMatrix<T>& V, const Faces& F, SparseMatrix<T>& L) {
    std::vector<Triplet<T>> triplets;
    triplets.reserve(F.size() * 12);
    for (const auto& f : F) {
        const Vec3<T> e0 = row(V, f[0]) - row(V, f[1]);
        const Vec3<T> e1 = row(V, f[0]) - row(V, f[2]);
        const T a = norm(e0.cross(e1));
        for (int j = 0; j < 3; ++j) {
            const int j1 = f[(j + 1) % 3];
            const int j2 = f[(j + 2) % 3];
            const Vec3<T> d0 = row(V, f[j]) - row(V, j1);
            const Vec3<T> d1 = row(V, f[j]) - row(V, j2);
            const T w = d0.dot(d1) / a;
            triplets.emplace_back(j1, j1, -w);
            triplets.emplace_back(j2, j2, -w);
            triplets.emplace_back(j1, j2, w);
            triplets.emplace_back(j2, j1, w);
        }
    }
    L.resize(V.rows(), V.rows());
    L.setFromTriplets(triplets.begin(), triplets.end());
}

/* Baraff-Witkin stretch + shear ***********************************************/

struct RestShape {
    std::vector<double> area;                 // a(f)
    std::vector<std::array<double, 4>> Vinv;  // V_inverse(f), row-major 2x2
};

// Triangle area and inverse reference frame per face (stretch_shear_forces_impl.hpp:23-49).
inline RestShape precomputeRestShape(const DenseMatrix<double>& V, const Faces& F) {
    RestShape r;
    r.area.resize(F.size());
    r.Vinv.resize(F.size());
    for (size_t i = 0; i < F.size(); ++i) {
        const Vec3<double> v1 = row(V, F[i][1]) - row(V, F[i][0]);
        const Vec3<double> v2 = row(V, F[i][2]) - row(V, F[i][0]);
        r.area[i] = 0.5 * norm(v1.cross(v2));
        const double l1 = norm(v1);
        const Vec3<double> u = {v1.x / l1, v1.y / l1, v1.z / l1};
        const Vec3<double> normal = u.cross(v2);
        Vec3<double> w = normal.cross(u);
        const double lw = norm(w);
        w = {w.x / lw, w.y / lw, w.z / lw};
        // Vref = [[|v1|, v2.u], [0, v2.w]]; inverse by cofactors
        const double a = l1, b = v2.dot(u), c = 0.0, d = v2.dot(w);
        const double invdet = 1.0 / (a * d - b * c);
        r.Vinv[i] = {d * invdet, -b * invdet, -c * invdet, a * invdet};
    }
    return r;
}

// Per-face energy -k_stretch |C_stretch|^2 - k_shear C_shear^2 (stretch_shear_forces_impl.hpp:51-73).
template <class T>
std::vector<T> stretchShearEnergy(const DenseMatrix<T>& X, const Faces& F, const RestShape& r, double kStretch,
                                  double kShear) {
    std::vector<T> ret;
    ret.reserve(F.size());
    for (size_t f = 0; f < F.size(); ++f) {
        const Vec3<T> D0 = row(X, F[f][1]) - row(X, F[f][0]);
        const Vec3<T> D1 = row(X, F[f][2]) - row(X, F[f][0]);
        const auto& Vi = r.Vinv[f];
        const T v00 = T(Vi[0]), v01 = T(Vi[1]), v10 = T(Vi[2]), v11 = T(Vi[3]);
        const Vec3<T> J0 = scaled(D0, v00) + scaled(D1, v10);
        const Vec3<T> J1 = scaled(D0, v01) + scaled(D1, v11);
        const T a = T(r.area[f]);
        const T one = T(1.0);
        const T c0 = a * (norm(J0) - one);
        const T c1 = a * (norm(J1) - one);
        const T cs = a * J0.dot(J1);
        ret.push_back(T(-kStretch) * (c0 * c0 + c1 * c1) - T(kShear) * cs * cs);
    }
    return ret;
}

/* quadratic bending (constant Hessian) ****************************************/

// Four vertices (v0, v1 on the edge; v2, v3 opposite) of every interior edge
// (examples/BaraffWitkin/adjacency.cpp:131-203).  Edges are enumerated in increasing
// (min, max) vertex order.
inline std::vector<std::array<int, 4>> interiorEdgeQuads(const Faces& F) {
    std::map<std::pair<int, int>, std::vector<int>> edgeFaces;
    for (int f = 0; f < (int)F.size(); ++f)
        for (int k = 0; k < 3; ++k) {
            int a = F[f][k], b = F[f][(k + 1) % 3];
            edgeFaces[{std::min(a, b), std::max(a, b)}].push_back(f);
        }
    std::vector<std::array<int, 4>> E4;
    for (auto& e : edgeFaces) {
        if (e.second.size() != 2) continue;
        const int fa = e.second[0], fb = e.second[1];
        // orient (v0, v1) as in face fa, v2 = third vertex of fa, v3 = third vertex of fb
        int v0 = -1, v1 = -1, v2 = -1, v3 = -1;
        for (int k = 0; k < 3; ++k) {
            const int a = F[fa][k], b = F[fa][(k + 1) % 3];
            if (std::min(a, b) == e.first.first && std::max(a, b) == e.first.second) {
                v0 = a;
                v1 = b;
                v2 = F[fa][(k + 2) % 3];
            }
        }
        for (int k = 0; k < 3; ++k)
            if (F[fb][k] != v0 && F[fb][k] != v1) v3 = F[fb][k];
        E4.push_back({v0, v1, v2, v3});
    }
    return E4;
}

inline double cotTheta(const Vec3<double>& v, const Vec3<double>& w) { return v.dot(w) / norm(v.cross(w)); }

// 4x4 stiffness Q(e) = -3 / (A0 + A1) * K0 K0^T (bend_quadratic_forces.cpp:54-83), row-major.
inline std::array<double, 16> bendStiffness(const DenseMatrix<double>& X, const std::array<int, 4>& q) {
    const Vec3<double> x0 = row(X, q[0]), x1 = row(X, q[1]), x2 = row(X, q[2]), x3 = row(X, q[3]);
    const Vec3<double> e0 = x1 - x0, e1 = x2 - x0, e2 = x3 - x0, e3 = x2 - x1, e4 = x3 - x1;
    const Vec3<double> me0 = {-e0.x, -e0.y, -e0.z};
    const double c01 = cotTheta(e0, e1), c02 = cotTheta(e0, e2), c03 = cotTheta(me0, e3), c04 = cotTheta(me0, e4);
    const double K0[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
    const double A0 = norm(e0.cross(e1)), A1 = norm(e0.cross(e2));
    const double coef = -3. / (A0 + A1);
    std::array<double, 16> Q;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Q[i * 4 + j] = coef * K0[i] * K0[j];
    return Q;
}

// Triplets of k_bend * Q(e) on every coordinate block, as Bend::precompute_rest_shape
// (bend_quadratic_forces.cpp:110-155) but with the Hessian value in the caller's scalar type.
template <class T>
void appendBendTriplets(const DenseMatrix<double>& rest, const Faces& F, double kBend, int n,
                        std::vector<Triplet<T>>& out) {
    for (const auto& q : interiorEdgeQuads(F)) {
        const auto Q = bendStiffness(rest, q);
        for (int dim = 0; dim < 3; ++dim)
            for (int v1 = 0; v1 < 4; ++v1) {
                for (int v2 = 0; v2 < v1; ++v2) {
                    out.emplace_back(q[v1] + dim * n, q[v2] + dim * n, T(kBend * Q[v1 * 4 + v2]));
                    out.emplace_back(q[v2] + dim * n, q[v1] + dim * n, T(kBend * Q[v2 * 4 + v1]));
                }
                out.emplace_back(q[v1] + dim * n, q[v1] + dim * n, T(kBend * Q[v1 * 4 + v1]));
            }
    }
}

// Symbolic Hessian of the stretch + shear energy (AD, examples/BaraffWitkinAD/main.cpp:90-98)
// plus the constant bending stiffness.  XS: n x 3 matrix of variables {v + c n, 0}.
inline SparseMatrix<Symbolic> baraffWitkinHessian(const DenseMatrix<Symbolic>& XS, const DenseMatrix<double>& rest,
                                                  const Faces& F, double kStretch, double kShear, double kBend,
                                                  bool shareTemplates = true) {
    const int n = XS.rows();
    const RestShape r = precomputeRestShape(rest, F);
    Symbolic en(ADD, stretchShearEnergy(XS, F, r, kStretch, kShear));
    auto tuples = hessianSparseVars(en, std::vector<Symbolic>(XS.data(), XS.data() + XS.size()), shareTemplates);

    std::vector<Triplet<Symbolic>> triplets;
    triplets.reserve(2 * tuples.size() + 48 * F.size());
    for (const auto& t : tuples) triplets.emplace_back(std::get<0>(t), std::get<1>(t), std::get<2>(t));
    for (const auto& t : tuples)
        if (std::get<1>(t) != std::get<0>(t)) triplets.emplace_back(std::get<1>(t), std::get<0>(t), std::get<2>(t));
    if (kBend != 0.0) appendBendTriplets(rest, F, kBend, n, triplets);

    SparseMatrix<Symbolic> H(3 * n, 3 * n);
    H.setFromTriplets(triplets.begin(), triplets.end());
    return H;
}

}  // namespace workloads
}  // namespace Sym
