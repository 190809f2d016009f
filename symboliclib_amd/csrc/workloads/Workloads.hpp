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
#include <limits>
#include <map>
#include <stdexcept>
#include <vector>

#include "../sym/Assembly.hpp"
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

// The same Laplacian as one instance of a direct assembly (no per-nonzero expression): one
// element per face whose symmetric 3 x 3 element matrix holds the face's corner weights,
// M(j1, j2) = M(j2, j1) = w_j for the corner j opposite edge (j1, j2) and
// M(a, a) = -(w_b + w_c) over the two other corners, so that summing the element matrices adds
// exactly the twelve triplets of buildCotan per face.  Inputs: the n x 3 position matrix
// {v + d n, group}; output: the n x n Laplacian over the vertices (a node pattern).
class CotanShape {
    int id_ = -1;

public:
    int get(HessianAssembly& A) {
        if (id_ >= 0) return id_;
        // local a + 3 d: coordinate d of the face's vertex a, so locals 0..2 are the rows
        auto P = [](int a) {
            return Vec3<Symbolic>{Symbolic(a, ELEMENT_LOCAL), Symbolic(a + 3, ELEMENT_LOCAL), Symbolic(a + 6, ELEMENT_LOCAL)};
        };
        const Vec3<Symbolic> e0 = P(0) - P(1), e1 = P(0) - P(2);
        const Symbolic area2 = norm(e0.cross(e1));
        Symbolic w[3];
        for (int j = 0; j < 3; ++j) w[j] = (P(j) - P((j + 1) % 3)).dot(P(j) - P((j + 2) % 3)) / area2;
        Symbolic M[9];
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            M[j1 * 3 + j2] = M[j2 * 3 + j1] = w[j];
            M[j * 3 + j] = -(w[j1] + w[j2]);
        }
        const int rows[3] = {0, 1, 2};
        id_ = A.defineMatrixShape(3, M, rows, 9, 0);
        return id_;
    }
};

inline void cotanAssembly(HessianAssembly& A, CotanShape& shape, int n, const Faces& F, int group) {
    A.addInstance(3 * n, group, 3);
    A.setNodePattern();
    const int s = shape.get(A);
    int vars[9];
    for (const auto& f : F) {
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) vars[a + 3 * d] = f[a] + d * n;
        A.addElement(s, vars, nullptr);
    }
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

// Energy -k_stretch |C_stretch|^2 - k_shear C_shear^2 of face f (stretch_shear_forces_impl.hpp:51-73).
template <class T>
T faceEnergy(const DenseMatrix<T>& X, const std::array<int, 3>& fv, const RestShape& r, size_t f, double kStretch,
             double kShear) {
    const Vec3<T> D0 = row(X, fv[1]) - row(X, fv[0]);
    const Vec3<T> D1 = row(X, fv[2]) - row(X, fv[0]);
    const auto& Vi = r.Vinv[f];
    const T v00 = T(Vi[0]), v01 = T(Vi[1]), v10 = T(Vi[2]), v11 = T(Vi[3]);
    const Vec3<T> J0 = scaled(D0, v00) + scaled(D1, v10);
    const Vec3<T> J1 = scaled(D0, v01) + scaled(D1, v11);
    const T a = T(r.area[f]);
    const T one = T(1.0);
    const T c0 = a * (norm(J0) - one);
    const T c1 = a * (norm(J1) - one);
    const T cs = a * J0.dot(J1);
    return T(-kStretch) * (c0 * c0 + c1 * c1) - T(kShear) * cs * cs;
}

template <class T>
std::vector<T> stretchShearEnergy(const DenseMatrix<T>& X, const Faces& F, const RestShape& r, double kStretch,
                                  double kShear) {
    std::vector<T> ret;
    ret.reserve(F.size());
    for (size_t f = 0; f < F.size(); ++f) ret.push_back(faceEnergy(X, F[f], r, f, kStretch, kShear));
    return ret;
}

/* quadratic bending (constant Hessian) ****************************************/

// Four vertices (v0, v1 on the edge; v2, v3 opposite) of every interior edge
// (examples/BaraffWitkin/adjacency.cpp:131-203).  Edges are enumerated in increasing
// (min, max) vertex order.
inline std::vector<std::array<int, 4>> interiorEdgeQuads(const Faces& F) {
    // half-edges sorted by (min vertex, max vertex, face, corner)
    struct HalfEdge {
        int lo, hi, f, k;
    };
    std::vector<HalfEdge> he;
    he.reserve(3 * F.size());
    for (int f = 0; f < (int)F.size(); ++f)
        for (int k = 0; k < 3; ++k) {
            const int a = F[f][k], b = F[f][(k + 1) % 3];
            he.push_back({std::min(a, b), std::max(a, b), f, k});
        }
    std::sort(he.begin(), he.end(), [](const HalfEdge& x, const HalfEdge& y) {
        return x.lo != y.lo ? x.lo < y.lo : x.hi != y.hi ? x.hi < y.hi : x.f != y.f ? x.f < y.f : x.k < y.k;
    });
    std::vector<std::array<int, 4>> E4;
    for (size_t s = 0; s < he.size();) {
        size_t e = s + 1;
        while (e < he.size() && he[e].lo == he[s].lo && he[e].hi == he[s].hi) ++e;
        if (e - s == 2) {
            // orient (v0, v1) as in the first face, v2 = its third vertex, v3 = third vertex of the second
            const int fa = he[s].f, ka = he[s].k, fb = he[s + 1].f;
            const int v0 = F[fa][ka], v1 = F[fa][(ka + 1) % 3], v2 = F[fa][(ka + 2) % 3];
            int v3 = -1;
            for (int k = 0; k < 3; ++k)
                if (F[fb][k] != v0 && F[fb][k] != v1) v3 = F[fb][k];
            E4.push_back({v0, v1, v2, v3});
        }
        s = e;
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

// Element shapes of the face energy, one per pattern of zero rest constants (a zero stays a
// literal, so the terms it multiplies vanish from the derivatives as in the per-face trees; all
// other constants become parameters, so cloths of every size share the shapes).  Shared by all
// instances of an assembly.
class FaceShapes {
    std::map<std::array<int64_t, 5>, int> cache_;
    double kStretch_, kShear_;

public:
    FaceShapes(double kStretch, double kShear) : kStretch_(kStretch), kShear_(kShear) {}

    // shape of face f and its parameters (the non-integer constants among area, Vinv)
    int get(HessianAssembly& A, const RestShape& r, size_t f, double params[5]) {
        const double c[5] = {r.area[f], r.Vinv[f][0], r.Vinv[f][1], r.Vinv[f][2], r.Vinv[f][3]};
        std::array<int64_t, 5> key;
        for (int k = 0; k < 5; ++k) {
            // literal zero, or parameter (a code no literal takes)
            key[k] = c[k] == 0.0 ? 0 : std::numeric_limits<int64_t>::min();
            params[k] = c[k];
        }
        auto it = cache_.find(key);
        if (it != cache_.end()) return it->second;
        // build the energy once over local variables and parameters
        DenseMatrix<Symbolic> XL(3, 3);
        for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) XL(i, d) = Symbolic(3 * i + d, ELEMENT_LOCAL);
        RestShape rl;
        Symbolic sym[5];
        for (int k = 0; k < 5; ++k)
            sym[k] = key[k] == std::numeric_limits<int64_t>::min() ? Symbolic(k, ELEMENT_PARAM) : Symbolic(c[k]);
        const Vec3<Symbolic> D0 = row(XL, 1) - row(XL, 0);
        const Vec3<Symbolic> D1 = row(XL, 2) - row(XL, 0);
        const Vec3<Symbolic> J0 = scaled(D0, sym[1]) + scaled(D1, sym[3]);
        const Vec3<Symbolic> J1 = scaled(D0, sym[2]) + scaled(D1, sym[4]);
        const Symbolic one(1.0);
        const Symbolic c0 = sym[0] * (norm(J0) - one);
        const Symbolic c1 = sym[0] * (norm(J1) - one);
        const Symbolic cs = sym[0] * J0.dot(J1);
        const Symbolic e = Symbolic(-kStretch_) * (c0 * c0 + c1 * c1) - Symbolic(kShear_) * cs * cs;
        const int id = A.defineShape(e, 9, 5);
        cache_[key] = id;
        return id;
    }
};

// The same Hessian as one instance of a direct assembly; variables are {v + c n, group}.
// The face energy is the expression of faceEnergy() with the rest constants as parameters.
inline void baraffWitkinAssembly(HessianAssembly& A, FaceShapes& shapes, const DenseMatrix<double>& rest,
                                 const Faces& F, double kBend, int group) {
    const int n = rest.rows();
    A.addInstance(3 * n, group, 3);
    const RestShape r = precomputeRestShape(rest, F);
    double params[5];
    int vars[9];
    for (size_t f = 0; f < F.size(); ++f) {
        const int s = shapes.get(A, r, f, params);
        for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) vars[3 * i + d] = F[f][i] + d * n;
        A.addElement(s, vars, params);
    }
    if (kBend != 0.0)
        for (const auto& q : interiorEdgeQuads(F)) {
            // k_bend Q(e) on every coordinate block, as appendBendTriplets
            auto Q = bendStiffness(rest, q);
            for (double& v : Q) v *= kBend;
            A.addNodalConstant(4, q.data(), Q.data());
        }
}

}  // namespace workloads
}  // namespace Sym
