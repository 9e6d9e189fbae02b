// fem_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of SOFA's implicit corotational-FEM time step, used as the
// correctness oracle by tests/, by __graft_entry__.smoke() and as the CPU baseline leg
// of bench.py.  Nothing in the product (sofa_amd/) links, loads or calls this file.
//
// It follows the reference line by line, in the reference's own operation order:
//   TetrahedronFEMForceField<Vec3d>  Sofa/Component/SolidMechanics/FEM/Elastic/src/sofa/
//       component/solidmechanics/fem/elastic/TetrahedronFEMForceField.inl
//         computeStrainDisplacement :134-203, peudo_determinant_for_coef :205-209,
//         computeMaterialStiffness :256-291, computeForce (addForce form) :293-415,
//         computeForce (dForce form) :417-521, initSmall/accumulateForceSmall :526-620,
//         applyStiffnessSmall :724-749, computeRotationLarge :755-778, initLarge :835-868,
//         accumulateForceLarge :870-986, applyStiffnessCorotational :1193-1237,
//         init/reinit :1257-1545, addForce :1548-1604, addDForce :1606-1636
//   BaseLinearElasticityFEMForceField.inl:113-144 (per-element E / nu lookup)
//   sofa/type/Vec.h:483-580 (norm2 / normalize), :775-780 (cross); Mat.h:577-587 (Mat*Vec)
//   sofa/geometry/Tetrahedron.h:55-83 (volume = |signed volume|)
//   UniformMass.inl:329-345,403-419,471-496; DiagonalMass.inl:536-559,1060-1080,1393-1413
//   MeshMatrixMass.inl:547-657,1387-1494,1987-2092; TetrahedronSetTopologyContainer.cpp:105-147
//   FixedProjectiveConstraint.inl:184-207
//   OscillatorProjectiveConstraint.inl:65-106
//   CGLinearSolver.inl:71-315, CGLinearSolver.cpp:40-64 (GraphScattered cgstep_beta/alpha)
//   GraphScatteredTypes.cpp:33-45 (A*p = project((mM + bB + kK) p))
//   EulerImplicitSolver.cpp:83-300
//   MechanicalObject.inl:2076-2356 (vOp, vMultiOp, vDot), BaseMechanicalState.cpp:42-80
//   BaseForceField.cpp:38-47, Mass.inl:94-104, MechanicalParams.h:62-64
//   DefaultAnimationLoop.cpp:273-329 (solve, then projectPositionAndVelocity, then time += dt)
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, so that no multiply-add is fused:
// the reference's x86-64 build does not fuse either).
//
// Threads (orc_set_threads, the CPU baseline on all cores): the element loops of addForce and
// addDForce (method "large") run in parallel writing each element's 12 nodal values, then each
// node applies its values in ascending element id -- the same operations in the same order per
// node as the serial loop, so the result is bit-identical for any thread count.
//
// Dot-product order: the reference sums r.r node by node in ascending order
// (MechanicalObject.inl:2334-2356).  The GPU sums per 256-node block with a fixed tree and
// then over blocks.  `dot_mode = 1` replays that GPU order exactly (see dot_blocked), so the
// GPU trajectory can be compared bit for bit as well as against the reference order.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>
#include <array>
#include <memory>
#include <algorithm>
#include <map>

namespace oracle {

typedef double Real;

struct Vec3 {
    Real v[3];
    Vec3() { v[0] = v[1] = v[2] = 0; }
    Vec3(Real a, Real b, Real c) { v[0] = a; v[1] = b; v[2] = c; }
    Real& operator[](int i) { return v[i]; }
    Real operator[](int i) const { return v[i]; }
};

static inline Vec3 sub(const Vec3& a, const Vec3& b) { return Vec3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
// Vec.h:407-413 -- r = a0*b0; r += a1*b1; r += a2*b2
static inline Real dot3(const Vec3& a, const Vec3& b) {
    Real r = a[0] * b[0];
    r += a[1] * b[1];
    r += a[2] * b[2];
    return r;
}
// Vec.h:775-780
static inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return Vec3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
// Vec.h:483-492 (norm2), 497-500 (norm), 535-578 (normalizeWithNorm / normalized)
static inline Vec3 normalized(const Vec3& a) {
    Vec3 r = a;
    Real n2 = r[0] * r[0];
    n2 += r[1] * r[1];
    n2 += r[2] * r[2];
    Real n = std::sqrt(n2);
    if (n > std::numeric_limits<Real>::epsilon()) {
        for (int i = 0; i < 3; i++) r[i] /= n;
    }
    return r;
}

struct Mat3 {
    Real m[3][3];
    Mat3() { std::memset(m, 0, sizeof(m)); }
};
// Mat.h:577-587: r(i) = m(i,0)*v0; r(i) += m(i,j)*v[j]
static inline Vec3 mul(const Mat3& M, const Vec3& v) {
    Vec3 r;
    for (int i = 0; i < 3; i++) {
        r[i] = M.m[i][0] * v[0];
        r[i] += M.m[i][1] * v[1];
        r[i] += M.m[i][2] * v[2];
    }
    return r;
}
static inline Mat3 transpose(const Mat3& M) {
    Mat3 t;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t.m[i][j] = M.m[j][i];
    return t;
}

// Tetrahedron.h:55-83
static inline Real tet_volume(const Vec3& n0, const Vec3& n1, const Vec3& n2, const Vec3& n3) {
    Vec3 a = sub(n1, n0), b = sub(n2, n0), c = sub(n3, n0);
    return std::abs(dot3(cross(a, b), c) / static_cast<Real>(6));
}

typedef std::array<int, 4> Tetra;

// ------------------------------------------------------------------------------------------
// Force-field interface (ForceField.h:87,110; Mass.h:69-77)
// ------------------------------------------------------------------------------------------
struct ForceField {
    virtual ~ForceField() {}
    virtual void addForce(std::vector<Vec3>& f, const std::vector<Vec3>& x, const std::vector<Vec3>& v) = 0;
    // BaseForceField::addMBKdx (BaseForceField.cpp:38-47) / Mass::addMBKdx (Mass.inl:94-104)
    virtual void addMBKdx(std::vector<Vec3>& df, const std::vector<Vec3>& dx, Real mFact, Real bFact, Real kFact) = 0;
};

// ------------------------------------------------------------------------------------------
// TetrahedronFEMForceField<Vec3d>
// ------------------------------------------------------------------------------------------
struct TetraFEM : public ForceField {
    enum Method { LARGE = 0, SMALL = 1 };
    int method = LARGE;
    std::vector<Tetra> elems;
    std::vector<Vec3> initialPoints;
    std::vector<Real> young, poisson;
    Real rayleighStiffness = 0;      // BaseForceField.cpp:34 default 0

    // per-element data, reference layout (TetrahedronFEMForceField.h:111-128)
    std::vector<std::array<Real, 36>> K;       // MaterialStiffness 6x6
    std::vector<std::array<Real, 72>> J;       // StrainDisplacement 12x6
    std::vector<Mat3> rotations, initialRotations;
    std::vector<std::array<Vec3, 4>> rotatedInitial;
    Real restVolume = 0;

    // parallel element loops: node -> (element * 4 + vertex) in ascending element id, and the
    // per-element nodal values
    int nthreads = 1;
    std::vector<int> incPtr, incIdx;
    std::vector<Real> contrib;
    void ensureIncidence(size_t nNodes) {
        if (incPtr.size() == nNodes + 1) return;
        incPtr.assign(nNodes + 1, 0);
        for (const Tetra& t : elems)
            for (int k = 0; k < 4; k++) incPtr[t[k] + 1]++;
        for (size_t i = 0; i < nNodes; i++) incPtr[i + 1] += incPtr[i];
        incIdx.assign(4 * elems.size(), 0);
        std::vector<int> fill(incPtr.begin(), incPtr.end() - 1);
        for (size_t e = 0; e < elems.size(); e++)
            for (int k = 0; k < 4; k++) incIdx[fill[elems[e][k]]++] = (int)(4 * e + k);
        contrib.assign(12 * elems.size(), 0.0);
    }

    // computeVonMisesStress state (TetrahedronFEMForceField.h: elemLambda, elemMu, elemShapeFun)
    std::vector<Real> elemLambda, elemMu;
    std::vector<std::array<Real, 16>> elemShapeFun;

    // sofa/type/Mat.h:1104-1168, the general inverse: Gauss-Jordan with full pivoting (first
    // strictly largest |m1(i,j)| over unused rows i and columns j, scanned row by row), the pivot
    // row divided by the pivot, every other row reduced by it, rows permuted back at the end
    static bool invertMatrix4(const Real from[16], Real dest[16]) {
        Real m1[4][4], m2[4][4];
        int r[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0}, row[4] = {0, 0, 0, 0}, col[4] = {0, 0, 0, 0};
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                m1[i][j] = from[4 * i + j];
                m2[i][j] = (i == j) ? 1.0 : 0.0;
            }
        for (int k = 0; k < 4; k++) {
            Real pivot = 0;
            for (int i = 0; i < 4; i++) {
                if (row[i]) continue;
                for (int j = 0; j < 4; j++) {
                    if (col[j]) continue;
                    Real t = m1[i][j];
                    if (t < 0) t = -t;
                    if (t > pivot) {
                        pivot = t;
                        r[k] = i;
                        c[k] = j;
                    }
                }
            }
            if (std::fabs(pivot) <= std::numeric_limits<Real>::epsilon()) return false;   // equalsZero
            row[r[k]] = col[c[k]] = 1;
            pivot = m1[r[k]][c[k]];
            for (int j = 0; j < 4; j++) m1[r[k]][j] /= pivot;
            m1[r[k]][c[k]] = 1;
            for (int j = 0; j < 4; j++) m2[r[k]][j] /= pivot;
            for (int i = 0; i < 4; i++) {
                if (i == r[k]) continue;
                const Real f = m1[i][c[k]];
                Real t1[4], t2[4];
                for (int j = 0; j < 4; j++) t1[j] = m1[r[k]][j] * f;
                for (int j = 0; j < 4; j++) m1[i][j] -= t1[j];
                m1[i][c[k]] = 0;
                for (int j = 0; j < 4; j++) t2[j] = m2[r[k]][j] * f;
                for (int j = 0; j < 4; j++) m2[i][j] -= t2[j];
            }
        }
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                if (c[j] == i) row[i] = r[j];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) dest[4 * i + j] = m2[row[i]][j];
        return true;
    }

    // reinit's von Mises set-up (:1521-1542): matVert rows (1, X0[vertex k])
    bool initVonMises() {
        elemShapeFun.assign(elems.size(), {});
        for (size_t i = 0; i < elems.size(); i++) {
            Real mv[16];
            for (int k = 0; k < 4; k++) {
                mv[4 * k] = 1.0;
                for (int l = 1; l < 4; l++) mv[4 * k + l] = initialPoints[elems[i][k]][l - 1];
            }
            if (!invertMatrix4(mv, elemShapeFun[i].data())) return false;
        }
        return true;
    }

    // computeVonMisesStress (:2197-2347) at positions X; mode 1 corotational strain (method
    // large), 2 full Green strain; vME per element, vMN per node (mean over the node's tetrahedra)
    void computeVonMisesStress(int mode, const std::vector<Vec3>& X, std::vector<Real>& vME, std::vector<Real>& vMN) {
        const size_t N = X.size();
        std::vector<Vec3> U(N);
        for (size_t i = 0; i < initialPoints.size(); i++) U[i] = sub(X[i], initialPoints[i]);
        vME.assign(elems.size(), 0.0);
        for (size_t el = 0; el < elems.size(); el++) {
            const Tetra& idx = elems[el];
            const auto& shf = elemShapeFun[el];
            Real vStrain[6];
            Real gradU[3][3];
            if (mode == 2) {
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) {
                        gradU[k][l] = 0.0;
                        for (int m = 0; m < 4; m++) gradU[k][l] += shf[4 * (l + 1) + m] * U[idx[m]][k];
                    }
                // 0.5 (gradU + gradU^T + gradU^T gradU): Mat sums, the 3x3 product (Mat.h:1444-1470)
                Real strain[3][3];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        const Real sum = gradU[i][j] + gradU[j][i];
                        const Real prod = gradU[0][i] * gradU[0][j] + gradU[1][i] * gradU[1][j] + gradU[2][i] * gradU[2][j];
                        strain[i][j] = (Real)0.5 * (sum + prod);
                    }
                for (int i = 0; i < 3; i++) vStrain[i] = strain[i][i];
                vStrain[3] = strain[1][2];
                vStrain[4] = strain[0][2];
                vStrain[5] = strain[0][1];
            } else {
                // method large: the element frame of X, rotations[el] updated as a side effect
                const Mat3 R02 = computeRotationLarge(X, idx[0], idx[1], idx[2]);
                rotations[el] = transpose(R02);
                Vec3 deforme[4];
                for (int i = 0; i < 4; i++) deforme[i] = mul(R02, X[idx[i]]);
                deforme[1][0] -= deforme[0][0];
                deforme[2][0] -= deforme[0][0];
                deforme[2][1] -= deforme[0][1];
                deforme[3] = sub(deforme[3], deforme[0]);
                const auto& ri = rotatedInitial[el];
                Real D[12];
                D[0] = 0; D[1] = 0; D[2] = 0;
                D[3] = ri[1][0] - deforme[1][0];
                D[4] = 0; D[5] = 0;
                D[6] = ri[2][0] - deforme[2][0];
                D[7] = ri[2][1] - deforme[2][1];
                D[8] = 0;
                D[9] = ri[3][0] - deforme[3][0];
                D[10] = ri[3][1] - deforme[3][1];
                D[11] = ri[3][2] - deforme[3][2];
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) {
                        gradU[k][l] = 0.0;
                        for (int m = 0; m < 4; m++) gradU[k][l] += shf[4 * (l + 1) + m] * D[3 * m + k];
                    }
                Real strain[3][3];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) strain[i][j] = Real(0.5) * (gradU[i][j] + gradU[j][i]);
                for (int i = 0; i < 3; i++) vStrain[i] = strain[i][i];
                vStrain[3] = strain[1][2];
                vStrain[4] = strain[0][2];
                vStrain[5] = strain[0][1];
            }
            const Real lambda = elemLambda[el], mu = elemMu[el];
            Real sv[6];
            Real traceStrain = 0.0;
            for (int k = 0; k < 3; k++) {
                traceStrain += vStrain[k];
                sv[k] = vStrain[k] * 2 * mu;
            }
            for (int k = 3; k < 6; k++) sv[k] = vStrain[k] * 2 * mu;
            for (int k = 0; k < 3; k++) sv[k] += lambda * traceStrain;
            vME[el] = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2] - sv[0] * sv[1] - sv[1] * sv[2] - sv[2] * sv[0] +
                                3 * sv[3] * sv[3] + 3 * sv[4] * sv[4] + 3 * sv[5] * sv[5]);
            if (vME[el] < 1e-10) vME[el] = 0.0;
        }
        ensureIncidence(N);
        vMN.assign(N, 0.0);
        for (size_t dof = 0; dof < N; dof++) {
            const int s0 = incPtr[dof], s1 = incPtr[dof + 1];
            for (int q = s0; q < s1; q++) vMN[dof] += vME[incIdx[q] / 4];
            if (s1 > s0) vMN[dof] /= Real(s1 - s0);
        }
    }

    // the general Mat product (sofa/type/Mat.h:1427-1441): r = 0, then r(i,j) += m1(i,k) m2(k,j)
    // in the loop order i, k, j
    static void matmul(const Real* a, const Real* b, Real* r, int L, int C, int P) {
        for (int i = 0; i < L * P; i++) r[i] = 0;
        for (int i = 0; i < L; i++)
            for (int k = 0; k < C; k++) {
                const Real aik = a[i * C + k];
                for (int j = 0; j < P; j++) r[i * P + j] += aik * b[k * P + j];
            }
    }

    // computeStiffnessMatrix (:212-232): JKJt = J K J^T, S = RR JKJt, SR = S RR^T with RR the
    // 12x12 block diagonal of Rot
    void computeStiffnessMatrix(size_t i, const Mat3& Rot, Real SR[144]) {
        Real Jt[72], JKm[72], JKJt[144], RR[144], RRt[144], Sm[144];
        for (int r = 0; r < 12; r++)
            for (int c = 0; c < 6; c++) Jt[c * 12 + r] = J[i][r * 6 + c];
        matmul(J[i].data(), K[i].data(), JKm, 12, 6, 6);
        matmul(JKm, Jt, JKJt, 12, 6, 12);
        for (int q = 0; q < 144; q++) RR[q] = RRt[q] = 0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int blk = 0; blk < 4; blk++) {
                    RR[(a + 3 * blk) * 12 + b + 3 * blk] = Rot.m[a][b];
                    RRt[(a + 3 * blk) * 12 + b + 3 * blk] = Rot.m[b][a];
                }
        matmul(RR, JKJt, Sm, 12, 12, 12);
        matmul(Sm, RRt, SR, 12, 12, 12);
    }

    // addKToMatrix (:1921-1957) into a dense 3N x 3N matrix: block (n1, n2) of element IT adds
    // -tmp(3 n1 + i, 3 n2 + j) * k at the rows / columns of its nodes, elements in order
    void addKToMatrix(Real* mat, size_t nNodes, Real k) {
        const size_t W = 3 * nNodes;
        Mat3 ident;
        for (int a = 0; a < 3; a++) ident.m[a][a] = 1;
        Real tmp[144];
        for (size_t it = 0; it < elems.size(); it++) {
            computeStiffnessMatrix(it, method == SMALL ? ident : rotations[it], tmp);
            for (int n1 = 0; n1 < 4; n1++)
                for (int n2 = 0; n2 < 4; n2++)
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++)
                            mat[(3 * (size_t)elems[it][n1] + a) * W + 3 * (size_t)elems[it][n2] + b] +=
                                -tmp[(3 * n1 + a) * 12 + 3 * n2 + b] * k;
        }
    }

    Real& Kel(size_t i, int r, int c) { return K[i][r * 6 + c]; }
    Real& Jel(size_t i, int r, int c) { return J[i][r * 6 + c]; }

    // BaseLinearElasticityFEMForceField.inl:113-144
    Real vecRealInElement(size_t e, const std::vector<Real>& d, Real def) const {
        if (d.size() > e) return d[e];
        if (!d.empty()) return d[0];
        return def;
    }

    // :205-209  (M is 2x3, row-major m[r][c])
    static Real pd(const Real M[2][3]) {
        return M[0][1] * M[1][2] - M[1][1] * M[0][2] - M[0][0] * M[1][2] + M[1][0] * M[0][2] + M[0][0] * M[1][1] - M[1][0] * M[0][1];
    }

    // :134-203
    void computeStrainDisplacement(size_t i, const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d) {
        Real M[2][3];
        auto& Jm = J[i];
        std::fill(Jm.begin(), Jm.end(), 0.0);
        auto S = [&](int r, int cc, Real v) { Jm[r * 6 + cc] = v; };
        M[0][0] = b[1]; M[0][1] = c[1]; M[0][2] = d[1];
        M[1][0] = b[2]; M[1][1] = c[2]; M[1][2] = d[2];
        { Real v = -pd(M); S(0, 0, v); S(1, 3, v); S(2, 5, v); }
        M[0][0] = b[0]; M[0][1] = c[0]; M[0][2] = d[0];
        { Real v = pd(M); S(0, 3, v); S(1, 1, v); S(2, 4, v); }
        M[1][0] = b[1]; M[1][1] = c[1]; M[1][2] = d[1];
        { Real v = -pd(M); S(0, 5, v); S(1, 4, v); S(2, 2, v); }

        M[0][0] = c[1]; M[0][1] = d[1]; M[0][2] = a[1];
        M[1][0] = c[2]; M[1][1] = d[2]; M[1][2] = a[2];
        { Real v = pd(M); S(3, 0, v); S(4, 3, v); S(5, 5, v); }
        M[0][0] = c[0]; M[0][1] = d[0]; M[0][2] = a[0];
        { Real v = -pd(M); S(3, 3, v); S(4, 1, v); S(5, 4, v); }
        M[1][0] = c[1]; M[1][1] = d[1]; M[1][2] = a[1];
        { Real v = pd(M); S(3, 5, v); S(4, 4, v); S(5, 2, v); }

        M[0][0] = d[1]; M[0][1] = a[1]; M[0][2] = b[1];
        M[1][0] = d[2]; M[1][1] = a[2]; M[1][2] = b[2];
        { Real v = -pd(M); S(6, 0, v); S(7, 3, v); S(8, 5, v); }
        M[0][0] = d[0]; M[0][1] = a[0]; M[0][2] = b[0];
        { Real v = pd(M); S(6, 3, v); S(7, 1, v); S(8, 4, v); }
        M[1][0] = d[1]; M[1][1] = a[1]; M[1][2] = b[1];
        { Real v = -pd(M); S(6, 5, v); S(7, 4, v); S(8, 2, v); }

        M[0][0] = a[1]; M[0][1] = b[1]; M[0][2] = c[1];
        M[1][0] = a[2]; M[1][1] = b[2]; M[1][2] = c[2];
        { Real v = pd(M); S(9, 0, v); S(10, 3, v); S(11, 5, v); }
        M[0][0] = a[0]; M[0][1] = b[0]; M[0][2] = c[0];
        { Real v = -pd(M); S(9, 3, v); S(10, 1, v); S(11, 4, v); }
        M[1][0] = a[1]; M[1][1] = b[1]; M[1][2] = c[1];
        { Real v = pd(M); S(9, 5, v); S(10, 4, v); S(11, 2, v); }
    }

    // :256-291
    void computeMaterialStiffness(size_t i, int a, int b, int c, int d) {
        const Real youngModulus = 1.0f * vecRealInElement(i, young, 5000);   // localStiffnessFactor empty
        const Real poissonRatio = vecRealInElement(i, poisson, 0.45);
        auto& Km = K[i];
        std::fill(Km.begin(), Km.end(), 0.0);
        Km[0 * 6 + 0] = Km[1 * 6 + 1] = Km[2 * 6 + 2] = 1;
        Km[0 * 6 + 1] = Km[0 * 6 + 2] = Km[1 * 6 + 0] = Km[1 * 6 + 2] = Km[2 * 6 + 0] = Km[2 * 6 + 1] = poissonRatio / (1 - poissonRatio);
        Km[3 * 6 + 3] = Km[4 * 6 + 4] = Km[5 * 6 + 5] = (1 - 2 * poissonRatio) / (2 * (1 - poissonRatio));
        const Real s = (youngModulus * (1 - poissonRatio)) / ((1 + poissonRatio) * (1 - 2 * poissonRatio));
        for (auto& k : Km) k *= s;
        // :278-282 (computeVonMisesStress set): the Lame coefficients before the volume division
        elemLambda[i] = Km[0 * 6 + 1];
        elemMu[i] = Km[3 * 6 + 3];
        const Real vol = tet_volume(initialPoints[a], initialPoints[b], initialPoints[c], initialPoints[d]);
        restVolume += vol;
        const Real div = vol * 36;
        for (auto& k : Km) k /= div;
    }

    // :755-778
    static Mat3 computeRotationLarge(const std::vector<Vec3>& p, int a, int b, int c) {
        const Vec3 edgex = normalized(sub(p[b], p[a]));
        Vec3 edgey = sub(p[c], p[a]);
        const Vec3 edgez = normalized(cross(edgex, edgey));
        edgey = cross(edgez, edgex);
        Mat3 r;
        for (int k = 0; k < 3; k++) {
            r.m[0][k] = edgex[k];
            r.m[1][k] = edgey[k];
            r.m[2][k] = edgez[k];
        }
        return r;
    }

    // :835-868
    void initLarge(size_t i, int a, int b, int c, int d) {
        Mat3 R01 = computeRotationLarge(initialPoints, a, b, c);
        initialRotations[i] = transpose(R01);
        rotations[i] = initialRotations[i];
        auto& ri = rotatedInitial[i];
        ri[0] = mul(R01, initialPoints[a]);
        ri[1] = mul(R01, initialPoints[b]);
        ri[2] = mul(R01, initialPoints[c]);
        ri[3] = mul(R01, initialPoints[d]);
        for (int k = 1; k < 4; k++)
            for (int q = 0; q < 3; q++) ri[k][q] -= ri[0][q];
        ri[0] = Vec3(0, 0, 0);
        computeStrainDisplacement(i, ri[0], ri[1], ri[2], ri[3]);
    }

    // :526-531
    void initSmall(size_t i, int a, int b, int c, int d) {
        computeStrainDisplacement(i, initialPoints[a], initialPoints[b], initialPoints[c], initialPoints[d]);
    }

    // init (:1257-1378) + reinit (:1391-1545) for an explicit tetra list
    void init(const std::vector<Vec3>& rest, const std::vector<Tetra>& tets, const std::vector<Real>& E, const std::vector<Real>& nu, int meth) {
        method = meth;
        elems = tets;
        initialPoints = rest;
        young = E.empty() ? std::vector<Real>{5000} : E;
        poisson = nu.empty() ? std::vector<Real>{0.45} : nu;
        // BaseLinearElasticityFEMForceField::checkYoungModulus / checkPoissonRatio (:49-84)
        for (auto& y : young) if (y < 0) y = 5000;
        for (auto& p : poisson) if (p < 0 || p >= 0.5) p = 0.45;
        const size_t ne = elems.size();
        K.assign(ne, {});
        J.assign(ne, {});
        elemLambda.assign(ne, 0.0);
        elemMu.assign(ne, 0.0);
        elemShapeFun.clear();
        restVolume = 0;
        if (method == LARGE) {
            rotations.assign(ne, Mat3());
            initialRotations.assign(ne, Mat3());
            rotatedInitial.assign(ne, {});
        }
        for (size_t i = 0; i < ne; i++) {
            const Tetra& t = elems[i];
            computeMaterialStiffness(i, t[0], t[1], t[2], t[3]);
            if (method == LARGE) initLarge(i, t[0], t[1], t[2], t[3]);
            else initSmall(i, t[0], t[1], t[2], t[3]);
        }
    }

    // :293-415 (plasticity disabled: plasticMaxThreshold defaults to 0)
    void computeForceF(Real F[12], const Real D[12], size_t i) {
        auto Jf = [&](int r, int c) { return J[i][r * 6 + c]; };
        auto Kf = [&](int r, int c) { return K[i][r * 6 + c]; };
        Real JtD[6];
        JtD[0] = Jf(0, 0) * D[0] + Jf(3, 0) * D[3] + Jf(6, 0) * D[6] + Jf(9, 0) * D[9];
        JtD[1] = Jf(1, 1) * D[1] + Jf(4, 1) * D[4] + Jf(7, 1) * D[7] + Jf(10, 1) * D[10];
        JtD[2] = Jf(2, 2) * D[2] + Jf(5, 2) * D[5] + Jf(8, 2) * D[8] + Jf(11, 2) * D[11];
        JtD[3] = Jf(0, 3) * D[0] + Jf(1, 3) * D[1] + Jf(3, 3) * D[3] + Jf(4, 3) * D[4] + Jf(6, 3) * D[6] + Jf(7, 3) * D[7] + Jf(9, 3) * D[9] + Jf(10, 3) * D[10];
        JtD[4] = Jf(1, 4) * D[1] + Jf(2, 4) * D[2] + Jf(4, 4) * D[4] + Jf(5, 4) * D[5] + Jf(7, 4) * D[7] + Jf(8, 4) * D[8] + Jf(10, 4) * D[10] + Jf(11, 4) * D[11];
        JtD[5] = Jf(0, 5) * D[0] + Jf(2, 5) * D[2] + Jf(3, 5) * D[3] + Jf(5, 5) * D[5] + Jf(6, 5) * D[6] + Jf(8, 5) * D[8] + Jf(9, 5) * D[9] + Jf(11, 5) * D[11];
        Real KJtD[6];
        KJtD[0] = Kf(0, 0) * JtD[0] + Kf(0, 1) * JtD[1] + Kf(0, 2) * JtD[2];
        KJtD[1] = Kf(1, 0) * JtD[0] + Kf(1, 1) * JtD[1] + Kf(1, 2) * JtD[2];
        KJtD[2] = Kf(2, 0) * JtD[0] + Kf(2, 1) * JtD[1] + Kf(2, 2) * JtD[2];
        KJtD[3] = Kf(3, 3) * JtD[3];
        KJtD[4] = Kf(4, 4) * JtD[4];
        KJtD[5] = Kf(5, 5) * JtD[5];
        jkjtd(F, KJtD, i);
    }

    // the common tail of both computeForce forms: F = J * KJtD (:384-413 and :490-519)
    void jkjtd(Real F[12], const Real KJtD[6], size_t i) {
        auto Jf = [&](int r, int c) { return J[i][r * 6 + c]; };
        F[0] = Jf(0, 0) * KJtD[0] + Jf(0, 3) * KJtD[3] + Jf(0, 5) * KJtD[5];
        F[1] = Jf(1, 1) * KJtD[1] + Jf(1, 3) * KJtD[3] + Jf(1, 4) * KJtD[4];
        F[2] = Jf(2, 2) * KJtD[2] + Jf(2, 4) * KJtD[4] + Jf(2, 5) * KJtD[5];
        F[3] = Jf(3, 0) * KJtD[0] + Jf(3, 3) * KJtD[3] + Jf(3, 5) * KJtD[5];
        F[4] = Jf(4, 1) * KJtD[1] + Jf(4, 3) * KJtD[3] + Jf(4, 4) * KJtD[4];
        F[5] = Jf(5, 2) * KJtD[2] + Jf(5, 4) * KJtD[4] + Jf(5, 5) * KJtD[5];
        F[6] = Jf(6, 0) * KJtD[0] + Jf(6, 3) * KJtD[3] + Jf(6, 5) * KJtD[5];
        F[7] = Jf(7, 1) * KJtD[1] + Jf(7, 3) * KJtD[3] + Jf(7, 4) * KJtD[4];
        F[8] = Jf(8, 2) * KJtD[2] + Jf(8, 4) * KJtD[4] + Jf(8, 5) * KJtD[5];
        F[9] = Jf(9, 0) * KJtD[0] + Jf(9, 3) * KJtD[3] + Jf(9, 5) * KJtD[5];
        F[10] = Jf(10, 1) * KJtD[1] + Jf(10, 3) * KJtD[3] + Jf(10, 4) * KJtD[4];
        F[11] = Jf(11, 2) * KJtD[2] + Jf(11, 4) * KJtD[4] + Jf(11, 5) * KJtD[5];
    }

    // :417-521 (dForce form, with `KJtD *= fact`)
    void computeForceDF(Real F[12], const Real D[12], size_t i, Real fact) {
        auto Jf = [&](int r, int c) { return J[i][r * 6 + c]; };
        auto Kf = [&](int r, int c) { return K[i][r * 6 + c]; };
        Real JtD[6];
        JtD[0] = Jf(0, 0) * D[0] + Jf(3, 0) * D[3] + Jf(6, 0) * D[6] + Jf(9, 0) * D[9];
        JtD[1] = Jf(1, 1) * D[1] + Jf(4, 1) * D[4] + Jf(7, 1) * D[7] + Jf(10, 1) * D[10];
        JtD[2] = Jf(2, 2) * D[2] + Jf(5, 2) * D[5] + Jf(8, 2) * D[8] + Jf(11, 2) * D[11];
        JtD[3] = Jf(0, 3) * D[0] + Jf(1, 3) * D[1] + Jf(3, 3) * D[3] + Jf(4, 3) * D[4] + Jf(6, 3) * D[6] + Jf(7, 3) * D[7] + Jf(9, 3) * D[9] + Jf(10, 3) * D[10];
        JtD[4] = Jf(1, 4) * D[1] + Jf(2, 4) * D[2] + Jf(4, 4) * D[4] + Jf(5, 4) * D[5] + Jf(7, 4) * D[7] + Jf(8, 4) * D[8] + Jf(10, 4) * D[10] + Jf(11, 4) * D[11];
        JtD[5] = Jf(0, 5) * D[0] + Jf(2, 5) * D[2] + Jf(3, 5) * D[3] + Jf(5, 5) * D[5] + Jf(6, 5) * D[6] + Jf(8, 5) * D[8] + Jf(9, 5) * D[9] + Jf(11, 5) * D[11];
        Real KJtD[6];
        KJtD[0] = Kf(0, 0) * JtD[0] + Kf(0, 1) * JtD[1] + Kf(0, 2) * JtD[2];
        KJtD[1] = Kf(1, 0) * JtD[0] + Kf(1, 1) * JtD[1] + Kf(1, 2) * JtD[2];
        KJtD[2] = Kf(2, 0) * JtD[0] + Kf(2, 1) * JtD[1] + Kf(2, 2) * JtD[2];
        KJtD[3] = Kf(3, 3) * JtD[3];
        KJtD[4] = Kf(4, 4) * JtD[4];
        KJtD[5] = Kf(5, 5) * JtD[5];
        for (int k = 0; k < 6; k++) KJtD[k] *= static_cast<Real>(fact);
        jkjtd(F, KJtD, i);
    }

    // :870-986 (assembling == false, updateStiffnessMatrix == false)
    void accumulateForceLarge(std::vector<Vec3>& f, const std::vector<Vec3>& p, size_t i) {
        Real c[12];
        forceLarge(p, i, c);
        const Tetra& index = elems[i];
        for (int k = 0; k < 12; k += 3)
            for (int q = 0; q < 3; q++) f[index[k / 3]][q] += c[k + q];
    }
    // the element part of accumulateForceLarge: rotation, R F per vertex into out[12]
    void forceLarge(const std::vector<Vec3>& p, size_t i, Real out[12]) {
        const Tetra& index = elems[i];
        Mat3 R02 = computeRotationLarge(p, index[0], index[1], index[2]);
        rotations[i] = transpose(R02);
        Vec3 deforme[4];
        for (int k = 0; k < 4; k++) deforme[k] = mul(R02, p[index[k]]);
        deforme[1][0] -= deforme[0][0];
        deforme[2][0] -= deforme[0][0];
        deforme[2][1] -= deforme[0][1];
        for (int q = 0; q < 3; q++) deforme[3][q] -= deforme[0][q];
        const auto& ri = rotatedInitial[i];
        Real D[12];
        D[0] = 0; D[1] = 0; D[2] = 0;
        D[3] = ri[1][0] - deforme[1][0];
        D[4] = 0; D[5] = 0;
        D[6] = ri[2][0] - deforme[2][0];
        D[7] = ri[2][1] - deforme[2][1];
        D[8] = 0;
        D[9] = ri[3][0] - deforme[3][0];
        D[10] = ri[3][1] - deforme[3][1];
        D[11] = ri[3][2] - deforme[3][2];
        Real F[12];
        computeForceF(F, D, i);
        for (int k = 0; k < 12; k += 3) {
            Vec3 c = mul(rotations[i], Vec3(F[k], F[k + 1], F[k + 2]));
            for (int q = 0; q < 3; q++) out[k + q] = c[q];
        }
    }

    // :533-620 (assembling == false)
    void accumulateForceSmall(std::vector<Vec3>& f, const std::vector<Vec3>& p, size_t i) {
        const Tetra& t = elems[i];
        const int a = t[0], b = t[1], c = t[2], d = t[3];
        const auto& ip = initialPoints;
        Real D[12];
        D[0] = 0; D[1] = 0; D[2] = 0;
        for (int q = 0; q < 3; q++) {
            D[3 + q] = ip[b][q] - ip[a][q] - p[b][q] + p[a][q];
            D[6 + q] = ip[c][q] - ip[a][q] - p[c][q] + p[a][q];
            D[9 + q] = ip[d][q] - ip[a][q] - p[d][q] + p[a][q];
        }
        Real F[12];
        computeForceF(F, D, i);
        for (int q = 0; q < 3; q++) {
            f[a][q] += F[q];
            f[b][q] += F[3 + q];
            f[c][q] += F[6 + q];
            f[d][q] += F[9 + q];
        }
    }

    // :1548-1604
    void addForce(std::vector<Vec3>& f, const std::vector<Vec3>& x, const std::vector<Vec3>&) override {
        if (nthreads > 1 && method == LARGE) {
            ensureIncidence(f.size());
            const long E = (long)elems.size(), N = (long)f.size();
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (long i = 0; i < E; i++) forceLarge(x, (size_t)i, &contrib[12 * i]);
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (long nd = 0; nd < N; nd++)
                for (int s = incPtr[nd]; s < incPtr[nd + 1]; s++)
                    for (int q = 0; q < 3; q++) f[nd][q] += contrib[3 * incIdx[s] + q];
            return;
        }
        for (size_t i = 0; i < elems.size(); i++) {
            if (method == LARGE) accumulateForceLarge(f, x, i);
            else accumulateForceSmall(f, x, i);
        }
    }

    // :1193-1237
    void applyStiffnessCorotational(std::vector<Vec3>& f, const std::vector<Vec3>& x, size_t i, Real fact) {
        Real o[12];
        dforceCorotational(x, i, fact, o);
        const Tetra& t = elems[i];
        for (int k = 0; k < 4; k++)
            for (int q = 0; q < 3; q++) f[t[k]][q] -= o[3 * k + q];
    }
    // the element part of applyStiffnessCorotational: the value each vertex's df loses
    void dforceCorotational(const std::vector<Vec3>& x, size_t i, Real fact, Real out[12]) {
        const Tetra& t = elems[i];
        const Mat3& R = rotations[i];
        Real X[12];
        for (int k = 0; k < 4; k++) {
            const Vec3& u = x[t[k]];
            X[3 * k + 0] = R.m[0][0] * u[0] + R.m[1][0] * u[1] + R.m[2][0] * u[2];
            X[3 * k + 1] = R.m[0][1] * u[0] + R.m[1][1] * u[1] + R.m[2][1] * u[2];
            X[3 * k + 2] = R.m[0][2] * u[0] + R.m[1][2] * u[1] + R.m[2][2] * u[2];
        }
        Real F[12];
        computeForceDF(F, X, i, fact);
        for (int k = 0; k < 4; k++) {
            out[3 * k + 0] = R.m[0][0] * F[3 * k] + R.m[0][1] * F[3 * k + 1] + R.m[0][2] * F[3 * k + 2];
            out[3 * k + 1] = R.m[1][0] * F[3 * k] + R.m[1][1] * F[3 * k + 1] + R.m[1][2] * F[3 * k + 2];
            out[3 * k + 2] = R.m[2][0] * F[3 * k] + R.m[2][1] * F[3 * k + 1] + R.m[2][2] * F[3 * k + 2];
        }
    }

    // :724-749 (applyStiffnessSmall: X = dx, f -= F)
    void applyStiffnessSmall(std::vector<Vec3>& f, const std::vector<Vec3>& x, size_t i, Real fact) {
        const Tetra& t = elems[i];
        Real X[12];
        for (int k = 0; k < 4; k++)
            for (int q = 0; q < 3; q++) X[3 * k + q] = x[t[k]][q];
        Real F[12];
        computeForceDF(F, X, i, fact);
        for (int k = 0; k < 4; k++)
            for (int q = 0; q < 3; q++) f[t[k]][q] -= F[3 * k + q];
    }

    // :1606-1636; kFactor = (Real) kFactorIncludingRayleighDamping(mparams, rayleighStiffness)
    void addDForce(std::vector<Vec3>& df, const std::vector<Vec3>& dx, Real kFactor) {
        if (nthreads > 1 && method == LARGE) {
            ensureIncidence(df.size());
            const long E = (long)elems.size(), N = (long)df.size();
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (long i = 0; i < E; i++) dforceCorotational(dx, (size_t)i, kFactor, &contrib[12 * i]);
#pragma omp parallel for num_threads(nthreads) schedule(static)
            for (long nd = 0; nd < N; nd++)
                for (int s = incPtr[nd]; s < incPtr[nd + 1]; s++)
                    for (int q = 0; q < 3; q++) df[nd][q] -= contrib[3 * incIdx[s] + q];
            return;
        }
        for (size_t i = 0; i < elems.size(); i++) {
            if (method == SMALL) applyStiffnessSmall(df, dx, i, kFactor);
            else applyStiffnessCorotational(df, dx, i, kFactor);
        }
    }

    void addMBKdx(std::vector<Vec3>& df, const std::vector<Vec3>& dx, Real, Real bFact, Real kFact) override {
        // BaseForceField.cpp:38-47
        const Real kFactor = kFact + bFact * rayleighStiffness;
        if (kFactor != 0.0 || bFact != 0.0) addDForce(df, dx, kFactor);
    }
};

// ------------------------------------------------------------------------------------------
// Masses
// ------------------------------------------------------------------------------------------
struct MassBase : public ForceField {
    Vec3 gravity;
    Real rayleighMass = 0;   // Mass.h rayleighMass Data, default 0
    virtual void addMDx(std::vector<Vec3>& res, const std::vector<Vec3>& dx, Real factor) = 0;
    // Mass::accFromF: a = M^-1 f; false where the reference refuses (a consistent MeshMatrixMass)
    virtual bool accFromF(std::vector<Vec3>& a, const std::vector<Vec3>& f) = 0;
    // node i's entry on the diagonal of the assembled mass matrix (buildMassMatrix)
    virtual Real diagMass(size_t i) const = 0;
    void addMBKdx(std::vector<Vec3>& df, const std::vector<Vec3>& dx, Real mFact, Real bFact, Real) override {
        // Mass.inl:94-104: ForceField::addMBKdx calls Mass::addDForce, a no-op; then addMDx
        const Real m = mFact - bFact * rayleighMass;
        if (m != 0.0) addMDx(df, dx, m);
    }
};

// UniformMass<Vec3d> (UniformMass.inl)
struct UniformMass : public MassBase {
    Real vertexMass = 1;
    void addMDx(std::vector<Vec3>& res, const std::vector<Vec3>& dx, Real factor) override {
        Real m = vertexMass;                             // :403-419
        if (factor != 1.0) m *= Real(factor);
        for (size_t i = 0; i < res.size(); i++)
            for (int q = 0; q < 3; q++) res[i][q] += dx[i][q] * m;
    }
    bool accFromF(std::vector<Vec3>& a, const std::vector<Vec3>& f) override {
        const Real m = vertexMass;                       // :423-436
        for (size_t i = 0; i < f.size(); i++)
            for (int q = 0; q < 3; q++) a[i][q] = f[i][q] / m;
        return true;
    }
    Real diagMass(size_t) const override { return vertexMass; }
    void addForce(std::vector<Vec3>& f, const std::vector<Vec3>&, const std::vector<Vec3>&) override {
        Vec3 mg(gravity[0] * vertexMass, gravity[1] * vertexMass, gravity[2] * vertexMass);   // :471-496
        for (size_t i = 0; i < f.size(); i++)
            for (int q = 0; q < 3; q++) f[i][q] += mg[q];
    }
};

// DiagonalMass<Vec3d> with a tetrahedral topology (DiagonalMass.inl)
struct DiagonalMass : public MassBase {
    std::vector<Real> masses;
    // computeVertexMass (:1060-1080); with total = true, `value` is totalMass and the masses are
    // computed for density 1 then scaled by totalMass / their sum (initFromTotalMass, :1235-1257)
    void init(const std::vector<Vec3>& pos, const std::vector<Tetra>& tets, Real value, bool total = false) {
        masses.assign(pos.size(), 0.0);
        const Real density = total ? Real(1.0) : value;
        Real sum = 0;
        for (const Tetra& t : tets) {
            const Real vol = tet_volume(pos[t[0]], pos[t[1]], pos[t[2]], pos[t[3]]);
            const Real mass = (density * vol) / Real(4.0);
            for (int j = 0; j < 4; j++) {
                masses[t[j]] += mass;
                sum += mass;
            }
        }
        if (total) {
            const Real d = sum < std::numeric_limits<Real>::epsilon() ? Real(1.0) : Real(value / sum);
            for (Real& m : masses) m *= d;
        }
    }
    void addMDx(std::vector<Vec3>& res, const std::vector<Vec3>& dx, Real factor) override {
        const size_t n = masses.size();                  // :536-559
        if (factor == 1.0) {
            for (size_t i = 0; i < n; i++)
                for (int q = 0; q < 3; q++) res[i][q] += dx[i][q] * masses[i];
        } else {
            for (size_t i = 0; i < n; i++)
                for (int q = 0; q < 3; q++) res[i][q] += (dx[i][q] * masses[i]) * Real(factor);
        }
    }
    bool accFromF(std::vector<Vec3>& a, const std::vector<Vec3>& f) override {
        for (size_t i = 0; i < masses.size(); i++)      // :564-576
            for (int q = 0; q < 3; q++) a[i][q] = f[i][q] / masses[i];
        return true;
    }
    Real diagMass(size_t i) const override { return masses[i]; }
    void addForce(std::vector<Vec3>& f, const std::vector<Vec3>&, const std::vector<Vec3>&) override {
        for (size_t i = 0; i < masses.size(); i++)      // :1393-1413
            for (int q = 0; q < 3; q++) f[i][q] += gravity[q] * masses[i];
    }
};

// MeshMatrixMass<Vec3d> on a tetrahedral topology (MeshMatrixMass.inl): vertex mass rho V / 10
// per tet vertex (:547-600), edge mass rho V / 20 per tet edge (:603-650), edges enumerated as
// TetrahedronSetTopologyContainer::createEdgeSetArray does (TetrahedronSetTopologyContainer.cpp
// :105-147; local edges {0,1},{0,2},{0,3},{1,2},{1,3},{2,3}, Topology.cpp:44), lumping
// coefficient 2.5 for tetrahedra (:1476-1489)
struct MeshMatrixMass : public MassBase {
    bool lumping = false;
    Real lumpCoeff = 2.5;
    std::vector<Real> vertexMass, edgeMass;
    std::vector<std::array<int, 2>> edges;
    void init(const std::vector<Vec3>& pos, const std::vector<Tetra>& tets, const std::vector<Real>& density, bool lump) {
        lumping = lump;
        static const int le[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        std::map<std::pair<int, int>, int> emap;
        std::vector<std::array<int, 6>> edgesInTet(tets.size());
        edges.clear();
        for (size_t i = 0; i < tets.size(); i++)
            for (int j = 0; j < 6; j++) {
                int v1 = tets[i][le[j][0]], v2 = tets[i][le[j][1]];
                std::pair<int, int> e = v1 < v2 ? std::make_pair(v1, v2) : std::make_pair(v2, v1);
                auto it = emap.find(e);
                if (it == emap.end()) {
                    const int idx = (int)emap.size();
                    emap[e] = idx;
                    edges.push_back({e.first, e.second});
                    edgesInTet[i][j] = idx;
                } else {
                    edgesInTet[i][j] = it->second;
                }
            }
        vertexMass.assign(pos.size(), 0.0);
        edgeMass.assign(edges.size(), 0.0);
        auto dens = [&](size_t i) { return density.size() > i ? density[i] : density[0]; };
        if (!lumping) {
            for (size_t i = 0; i < tets.size(); i++) {
                const Tetra& t = tets[i];
                const Real vol = tet_volume(pos[t[0]], pos[t[1]], pos[t[2]], pos[t[3]]);
                const Real mass = (dens(i) * vol) / Real(20.0);
                for (int j = 0; j < 6; j++) edgeMass[edgesInTet[i][j]] += mass;
            }
        }
        for (size_t i = 0; i < tets.size(); i++) {
            const Tetra& t = tets[i];
            const Real vol = tet_volume(pos[t[0]], pos[t[1]], pos[t[2]], pos[t[3]]);
            const Real mass = (dens(i) * vol) / Real(10.0);
            for (int j = 0; j < 4; j++) vertexMass[t[j]] += mass;
        }
    }
    // :1987-2047
    void addMDx(std::vector<Vec3>& res, const std::vector<Vec3>& dx, Real factor) override {
        if (lumping) {
            for (size_t i = 0; i < dx.size(); i++)
                for (int k = 0; k < 3; k++) res[i][k] += ((dx[i][k] * vertexMass[i]) * lumpCoeff) * Real(factor);
        } else {
            for (size_t i = 0; i < dx.size(); i++)
                for (int k = 0; k < 3; k++) res[i][k] += (dx[i][k] * vertexMass[i]) * Real(factor);
            for (size_t j = 0; j < edges.size(); j++) {
                const Real tempMass = edgeMass[j] * Real(factor);
                const int e0 = edges[j][0], e1 = edges[j][1];
                for (int k = 0; k < 3; k++) res[e0][k] += dx[e1][k] * tempMass;
                for (int k = 0; k < 3; k++) res[e1][k] += dx[e0][k] * tempMass;
            }
        }
    }
    // buildMassMatrix: the lumped vertex mass, or the consistent matrix's vertex term
    Real diagMass(size_t i) const override { return lumping ? vertexMass[i] * lumpCoeff : vertexMass[i]; }
    // :2051-2069: lumped only (the reference refuses the sparse mass)
    bool accFromF(std::vector<Vec3>& a, const std::vector<Vec3>& f) override {
        if (!lumping) return false;
        for (size_t i = 0; i < vertexMass.size(); i++)
            for (int k = 0; k < 3; k++) a[i][k] = f[i][k] / (vertexMass[i] * lumpCoeff);
        return true;
    }
    // :2073-2092 (gravity always with the lumped vertex mass)
    void addForce(std::vector<Vec3>& f, const std::vector<Vec3>&, const std::vector<Vec3>&) override {
        for (size_t i = 0; i < f.size(); i++)
            for (int k = 0; k < 3; k++) f[i][k] += (gravity[k] * vertexMass[i]) * lumpCoeff;
    }
};

// Test-only 1-D linear spring toward an anchor (used to pin the Euler recurrence of
// EulerImplicitSolverDynamic_test.cpp:76-101 with the closed form): f_y = -K (x_y - y0).
struct AxisSpring : public ForceField {
    int node = 0, axis = 1;
    Real stiffness = 0, anchor = 0;
    void addForce(std::vector<Vec3>& f, const std::vector<Vec3>& x, const std::vector<Vec3>&) override {
        f[node][axis] += -stiffness * (x[node][axis] - anchor);
    }
    void addMBKdx(std::vector<Vec3>& df, const std::vector<Vec3>& dx, Real, Real, Real kFact) override {
        if (kFact != 0.0) df[node][axis] += -stiffness * dx[node][axis] * kFact;
    }
};

// PlaneForceField<Vec3d> (PlaneForceField.inl:158-224): a penalty force on the nodes below the
// plane x.n = d (both sides if bilateral), recorded as contacts for addDForce
struct PlaneForceField : public ForceField {
    Vec3 normal = Vec3(0, 1, 0);
    Real d = 0, stiffness = 500, damping = 5, maxForce = 0;
    bool bilateral = false;
    std::vector<int> contacts;
    void addForce(std::vector<Vec3>& f, const std::vector<Vec3>& x, const std::vector<Vec3>& v) override {
        contacts.clear();
        Real limit = maxForce;
        limit *= limit;
        for (size_t i = 0; i < x.size(); i++) {
            const Real dist = dot3(x[i], normal) - d;
            if (bilateral || dist < 0) {
                const Real forceIntensity = -stiffness * dist;
                const Real dampingIntensity = -damping * dist;
                Vec3 force;
                for (int k = 0; k < 3; k++) force[k] = normal[k] * forceIntensity - v[i][k] * dampingIntensity;
                const Real amplitude = dot3(force, force);
                if (limit > 0 && amplitude > limit) {
                    const Real sc = std::sqrt(limit / amplitude);
                    for (int k = 0; k < 3; k++) force[k] *= sc;
                }
                for (int k = 0; k < 3; k++) f[i][k] += force[k];
                contacts.push_back((int)i);
            }
        }
    }
    void addMBKdx(std::vector<Vec3>& df, const std::vector<Vec3>& dx, Real, Real bFact, Real kFact) override {
        // BaseForceField.cpp:38-47 (rayleighStiffness of the component: 0), then addDForce
        if (kFact == 0.0 && bFact == 0.0) return;
        const Real fact = -stiffness * kFact;
        for (int p : contacts) {
            const Real a = fact * dot3(dx[p], normal);
            for (int k = 0; k < 3; k++) df[p][k] = df[p][k] + normal[k] * a;
        }
    }
};

// ------------------------------------------------------------------------------------------
// Scene: MechanicalObject + mass + FEM (+ test springs) + projective constraints,
// EulerImplicitSolver + CGLinearSolver<GraphScatteredMatrix, GraphScatteredVector>
// ------------------------------------------------------------------------------------------
struct Oscillator { int index; Vec3 mean, amplitude; Real pulsation, phase; };

// LinearMovementProjectiveConstraint<Vec3d> (LinearMovementProjectiveConstraint.inl:40-330):
// indices, keyTimes, movements, relativeMovements; x0 captured at the first projectPosition
struct LinearMovement {
    std::vector<int> indices;
    std::vector<Real> keyTimes;
    std::vector<Vec3> movements;
    bool relative = true;
    std::vector<Vec3> x0;
    bool x0set = false;
    // findKeyTimes (:296-330) at context time cT: true when cT lies in [keyTimes[k],
    // keyTimes[k+1]) for some k ("finished" with nextT != prevT); prev / next the bracketing keys
    bool window(Real cT, Real& prevT, Real& nextT, Vec3& prevM, Vec3& nextM) const {
        bool finished = false;
        if (!keyTimes.empty() && cT >= keyTimes.front() && cT <= keyTimes.back()) {
            nextT = keyTimes.front();
            prevT = nextT;
            for (size_t k = 0; k < keyTimes.size() && !finished; k++) {
                if (keyTimes[k] <= cT) {
                    prevT = keyTimes[k];
                    prevM = movements[k];
                } else {
                    nextT = keyTimes[k];
                    nextM = movements[k];
                    finished = true;
                }
            }
        }
        return finished && nextT != prevT;
    }
    // projectResponseT (:163-183): clear the constrained responses inside a movement window
    void projectResponse(std::vector<Vec3>& r, Real cT) const {
        Real pT = 0, nT = 0;
        Vec3 pM, nM;
        if (!window(cT, pT, nT, pM, nM)) return;
        for (int i : indices) r[i] = Vec3();
    }
    // projectVelocity (:193-214): v = (nextM - prevM) * (1 / (nextT - prevT))
    void projectVelocity(std::vector<Vec3>& vv, Real cT) const {
        Real pT = 0, nT = 0;
        Vec3 pM, nM;
        if (!window(cT, pT, nT, pM, nM)) return;
        const Real inv = 1.0 / (nT - pT);
        for (int i : indices)
            for (int k = 0; k < 3; k++) vv[i][k] = (nM[k] - pM[k]) * inv;
    }
    // projectPosition (:218-244) + interpolatePosition (:248-270): x0 captured on the first call;
    // m = prevM + (nextM - prevM) * ((cT - prevT) / (nextT - prevT)); x = x0 + m (relative) or m
    void projectPosition(std::vector<Vec3>& xx, Real cT) {
        if (!x0set) {
            x0.assign(xx.size(), Vec3());
            for (int i : indices) x0[i] = xx[i];
            x0set = true;
        }
        Real pT = 0, nT = 0;
        Vec3 pM, nM;
        if (!window(cT, pT, nT, pM, nM)) return;
        const Real dt = (cT - pT) / (nT - pT);
        Vec3 m;
        for (int k = 0; k < 3; k++) m[k] = pM[k] + (nM[k] - pM[k]) * dt;
        for (int i : indices)
            for (int k = 0; k < 3; k++) xx[i][k] = relative ? x0[i][k] + m[k] : m[k];
    }
};

struct Scene {
    std::vector<Vec3> x, v, f, x0;
    std::unique_ptr<MassBase> mass;
    std::unique_ptr<TetraFEM> fem;
    std::vector<std::unique_ptr<ForceField>> extra;
    std::vector<int> fixed;
    bool fixAll = false;
    bool projectVelocity = false;   // FixedProjectiveConstraint activate_projectVelocity (default false)
    std::vector<Oscillator> oscillators;
    std::vector<LinearMovement> movements;     // after the oscillators in scene order
    // PCGLinearSolver (PCGLinearSolver.inl:178-262) in place of CGLinearSolver: 0 none, 1 with a
    // JacobiPreconditioner, 2 with a BlockJacobiPreconditioner; its solution vector persists
    // across steps (the EulerImplicitSolver's x, realloc'd once at init), so each solve starts
    // from the previous step's solution
    int precond = 0;
    // CGLinearSolver warmStart (CGLinearSolver.inl:39,94-105): the solve starts from the solution
    // vector's current value, the previous step's solution (the same persistent vector as PCG's)
    bool warmStart = false;
    std::vector<Vec3> xsol;
    int lastPcgIterations = 0;
    Vec3 gravity;
    Real rayleighStiffness = 0, rayleighMass = 0;
    // EulerImplicitSolver Data (EulerImplicitSolver.cpp:40-51): firstOrder, trapezoidalScheme,
    // vdamping (all off by default)
    bool firstOrder = false, trapezoidal = false;
    Real velocityDamping = 0;
    unsigned cgIterations = 25;
    Real cgTolerance = 1e-5, cgThreshold = 1e-5;
    int dotMode = 0;           // 0: reference order; 1: GPU block order (256-node blocks)
    Real time = 0;
    int timeStepCount = 0;     // CGLinearSolver::timeStepCount
    bool equilibriumReached = false;
    // statistics of the last solve
    int lastIterations = 0;
    int lastEndCond = 0;       // 0 iterations, 1 tolerance, 2 threshold, 3 den==0, 4 null b
    Real lastErr = 0, lastDen = 0;

    size_t n() const { return x.size(); }

    // force fields in scene order: mass first, then the FEM, then extras
    std::vector<ForceField*> forceFields() {
        std::vector<ForceField*> r;
        if (mass) r.push_back(mass.get());
        if (fem) r.push_back(fem.get());
        for (auto& e : extra) r.push_back(e.get());
        return r;
    }

    // MechanicalObject::vDot (MechanicalObject.inl:2334-2356): r += va[i]*vb[i]
    Real dot_ref(const std::vector<Vec3>& a, const std::vector<Vec3>& b) const {
        Real r = 0.0;
        for (size_t i = 0; i < a.size(); i++) r += dot3(a[i], b[i]);
        return r;
    }
    // GPU order: per node s_i = dot3 (same as above); 256-node blocks reduced by the tree
    // sh[t] += sh[t+s] for s = 128..1; block partials reduced by a 256-thread pass where thread
    // t sums partials t, t+256, ... from 0, followed by the same tree.
    Real dot_blocked(const std::vector<Vec3>& a, const std::vector<Vec3>& b) const {
        const size_t BS = 256;
        const size_t nb = (a.size() + BS - 1) / BS;
        std::vector<Real> partial(nb);
        Real sh[BS];
        for (size_t blk = 0; blk < nb; blk++) {
            for (size_t t = 0; t < BS; t++) {
                size_t i = blk * BS + t;
                sh[t] = (i < a.size()) ? dot3(a[i], b[i]) : 0.0;
            }
            for (size_t s = BS / 2; s > 0; s >>= 1)
                for (size_t t = 0; t < s; t++) sh[t] = sh[t] + sh[t + s];
            partial[blk] = sh[0];
        }
        for (size_t t = 0; t < BS; t++) {
            Real acc = 0.0;
            for (size_t k = t; k < nb; k += BS) acc = acc + partial[k];
            sh[t] = acc;
        }
        for (size_t s = BS / 2; s > 0; s >>= 1)
            for (size_t t = 0; t < s; t++) sh[t] = sh[t] + sh[t + s];
        return sh[0];
    }
    Real vdot(const std::vector<Vec3>& a, const std::vector<Vec3>& b) const {
        return dotMode == 1 ? dot_blocked(a, b) : dot_ref(a, b);
    }

    // FixedProjectiveConstraint::projectResponse (:184-207) then Oscillator (:65-75)
    void projectResponse(std::vector<Vec3>& r) {
        if (fixAll) { for (auto& e : r) e = Vec3(); }
        else for (int i : fixed) r[i] = Vec3();
        for (const auto& o : oscillators) r[o.index] = Vec3();
        for (const auto& lm : movements) lm.projectResponse(r, time);
    }

    // projectVelocity at context time t: FixedProjectiveConstraint (.inl:237-258, when
    // activate_projectVelocity) then OscillatorProjectiveConstraint (.inl:72-87)
    void projectVelocityAt(std::vector<Vec3>& vv, Real t) {
        if (projectVelocity) {
            if (fixAll) { for (auto& e : vv) e = Vec3(); }
            else for (int i : fixed) vv[i] = Vec3();
        }
        for (const auto& o : oscillators)
            for (int k = 0; k < 3; k++) vv[o.index][k] = o.amplitude[k] * o.pulsation * std::cos(o.pulsation * t + o.phase);
        for (const auto& lm : movements) lm.projectVelocity(vv, t);
    }
    // projectPosition at context time t: Fixed is a no-op (.inl:262-265); Oscillator (:89-106)
    void projectPositionAt(std::vector<Vec3>& xx, Real t) {
        for (const auto& o : oscillators)
            for (int k = 0; k < 3; k++) xx[o.index][k] = o.mean[k] + o.amplitude[k] * std::sin(o.pulsation * t + o.phase);
        for (auto& lm : movements) lm.projectPosition(xx, t);
    }

    // GraphScatteredMatrix::apply (GraphScatteredTypes.cpp:33-45)
    void applyA(std::vector<Vec3>& q, const std::vector<Vec3>& p, Real mF, Real bF, Real kF) {
        for (auto& e : q) e = Vec3();                         // propagateDxAndResetDf
        for (ForceField* ff : forceFields()) ff->addMBKdx(q, p, mF, bF, kF);
        projectResponse(q);
    }

    // The preconditioner's matrix as MatrixLinearSystem assembles it: mF M (buildMassMatrix's
    // diagonal) then kF (-R J K J^T R^T) of every element in order (buildStiffnessMatrix's node
    // blocks), each value scaled by its factor as it is added; the 3x3 diagonal blocks are kept
    // (block Jacobi: BlockDiagonalMatrix<3>; Jacobi: DiagonalMatrix keeps their diagonals).
    // Dirichlet rows (FixedProjectiveConstraint, LinearMovementProjectiveConstraint indices:
    // applyConstraint(ZeroDirichletCondition*), MatrixLinearSystem.inl:1177-1189) become the
    // identity.  Then inverted: DiagonalMatrix::invert (1 / d) or Mat<3,3>::invert
    // (invertMatrix, Mat.h:1172-1192, with determinant, Mat.h:989-997).
    void precondInverse(std::vector<Mat3>& P, Real mF, Real kF) {
        const size_t N = n();
        P.assign(N, Mat3());
        if (mass)
            for (size_t i = 0; i < N; i++) {
                const Real m = mass->diagMass(i) * mF;
                for (int a = 0; a < 3; a++) P[i].m[a][a] += m;
            }
        if (fem) {
            Mat3 ident;
            for (int a = 0; a < 3; a++) ident.m[a][a] = 1;
            Real tmp[144];
            for (size_t it = 0; it < fem->elems.size(); it++) {
                fem->computeStiffnessMatrix(it, fem->method == TetraFEM::SMALL ? ident : fem->rotations[it], tmp);
                for (int v = 0; v < 4; v++) {
                    Mat3& B = P[fem->elems[it][v]];
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) B.m[a][b] += (-tmp[(3 * v + a) * 12 + 3 * v + b]) * kF;
                }
            }
        }
        auto dirichlet = [&](size_t i) {
            P[i] = Mat3();
            for (int a = 0; a < 3; a++) P[i].m[a][a] = 1;
        };
        if (fixAll) for (size_t i = 0; i < N; i++) dirichlet(i);
        else for (int i : fixed) dirichlet(i);
        for (const auto& lm : movements)
            for (int i : lm.indices) dirichlet(i);
        for (size_t i = 0; i < N; i++) {
            Mat3& B = P[i];
            if (precond == 1) {
                for (int a = 0; a < 3; a++) {
                    B.m[a][a] = 1.0 / B.m[a][a];
                    for (int b = 0; b < 3; b++) if (a != b) B.m[a][b] = 0;
                }
            } else {
                const Mat3 f = B;
                auto F = [&](int r, int c) { return f.m[r][c]; };
                const Real det = F(0, 0) * F(1, 1) * F(2, 2) + F(1, 0) * F(2, 1) * F(0, 2) + F(2, 0) * F(0, 1) * F(1, 2)
                                 - F(0, 0) * F(2, 1) * F(1, 2) - F(1, 0) * F(0, 1) * F(2, 2) - F(2, 0) * F(1, 1) * F(0, 2);
                B.m[0][0] = (F(1, 1) * F(2, 2) - F(2, 1) * F(1, 2)) / det;
                B.m[1][0] = (F(1, 2) * F(2, 0) - F(2, 2) * F(1, 0)) / det;
                B.m[2][0] = (F(1, 0) * F(2, 1) - F(2, 0) * F(1, 1)) / det;
                B.m[0][1] = (F(2, 1) * F(0, 2) - F(0, 1) * F(2, 2)) / det;
                B.m[1][1] = (F(2, 2) * F(0, 0) - F(0, 2) * F(2, 0)) / det;
                B.m[2][1] = (F(2, 0) * F(0, 1) - F(0, 0) * F(2, 1)) / det;
                B.m[0][2] = (F(0, 1) * F(1, 2) - F(1, 1) * F(0, 2)) / det;
                B.m[1][2] = (F(0, 2) * F(1, 0) - F(1, 2) * F(0, 0)) / det;
                B.m[2][2] = (F(0, 0) * F(1, 1) - F(1, 0) * F(0, 1)) / det;
            }
        }
    }
    // z = P^-1 r: DiagonalMatrix::mult (z_i = d_i r_i, DiagonalMatrix.h:187-190) or
    // BlockDiagonalMatrix::mul (z_j = sum_i B(i, j) r_i from 0, BlockDiagonalMatrix.h:202-221)
    void applyPrecond(const std::vector<Mat3>& P, const std::vector<Vec3>& r, std::vector<Vec3>& z) const {
        for (size_t i = 0; i < r.size(); i++) {
            if (precond == 1) {
                for (int a = 0; a < 3; a++) z[i][a] = P[i].m[a][a] * r[i][a];
            } else {
                for (int j = 0; j < 3; j++) {
                    Real acc = 0;
                    for (int a = 0; a < 3; a++) acc += P[i].m[a][j] * r[i][a];
                    z[i][j] = acc;
                }
            }
        }
    }
    // PCGLinearSolver::solve (PCGLinearSolver.inl:178-262) with the Jacobi / block-Jacobi
    // preconditioner, from the solution vector's current value
    void pcgSolve(std::vector<Vec3>& x, const std::vector<Vec3>& b, Real mF, Real bF, Real kF) {
        const size_t N = n();
        std::vector<Mat3> P;
        precondInverse(P, mF, kF);
        std::vector<Vec3> r(N), w(N), s(N);
        const Real b_norm = vdot(b, b);
        const Real tol = cgTolerance * b_norm;
        applyA(r, x, mF, bF, kF);                                   // r = M * x
        for (size_t i = 0; i < N; i++)                              // r = -1 r + b (r *= -1; r += b)
            for (int k = 0; k < 3; k++) { r[i][k] *= -1.0; r[i][k] += b[i][k]; }
        applyPrecond(P, r, w);
        Real r_norm = vdot(r, w);
        unsigned iter = 1;
        lastErr = b_norm != 0 ? r_norm / b_norm : 0;
        lastDen = 0;
        while (iter <= cgIterations && r_norm > tol) {
            applyA(s, w, mF, bF, kF);                               // s = M * w
            const Real dtq = vdot(w, s);
            const Real alpha = r_norm / dtq;
            lastDen = dtq;
            for (size_t i = 0; i < N; i++)
                for (int k = 0; k < 3; k++) {
                    x[i][k] += w[i][k] * alpha;                     // x.peq(w, alpha)
                    r[i][k] += s[i][k] * (-alpha);                  // r.peq(s, -alpha)
                }
            applyPrecond(P, r, s);
            const Real deltaOld = r_norm;
            r_norm = vdot(r, s);
            lastErr = b_norm != 0 ? r_norm / b_norm : 0;
            const Real beta = r_norm / deltaOld;
            for (size_t i = 0; i < N; i++)                          // w = s + w beta (w *= beta; w += s)
                for (int k = 0; k < 3; k++) { w[i][k] *= beta; w[i][k] += s[i][k]; }
            iter++;
        }
        lastIterations = (int)iter;
        lastEndCond = iter > cgIterations ? 0 : 1;
        timeStepCount++;
    }

    // CGLinearSolver<GraphScattered>::solve (CGLinearSolver.inl:71-315)
    void cgSolve(std::vector<Vec3>& xs, const std::vector<Vec3>& b, Real mF, Real bF, Real kF, bool warm = false) {
        const size_t N = n();
        std::vector<Vec3> p(N), q(N), r(N);
        Real rho, rho_1 = 0, alpha, beta;
        if (warm) {
            applyA(r, xs, mF, bF, kF);                        // r = A * x
            for (size_t i = 0; i < N; i++)                    // r.eq(b, r, -1.0): r = b + r (-1)
                for (int k = 0; k < 3; k++) r[i][k] = b[i][k] + r[i][k] * (-1.0);
        } else {
            for (auto& e : xs) e = Vec3();                    // x.clear()
            r = b;                                            // r = b
        }
        const Real normb = std::sqrt(vdot(b, b));             // b.norm()
        unsigned nb_iter = 0;
        int endcond = 0;
        lastErr = 1; lastDen = 0;
        if (normb != 0.0) {
            for (nb_iter = 1; nb_iter <= cgIterations; nb_iter++) {
                rho = vdot(r, r);
                const Real normr = std::sqrt(rho);
                const Real err = normr / normb;
                lastErr = err;
                if (err <= cgTolerance) {
                    if (nb_iter == 1 && timeStepCount == 0) {
                        // warning only: the loop goes on
                    } else {
                        if (nb_iter == 1 && !equilibriumReached) equilibriumReached = true;
                        if (nb_iter > 1) equilibriumReached = false;
                        endcond = 1;
                        break;
                    }
                }
                if (nb_iter == 1) {
                    p = r;
                } else {
                    beta = rho / rho_1;
                    // cgstep_beta, GraphScattered: p.eq(r, p, beta) -> vOp_avf: p *= beta; p += r
                    for (size_t i = 0; i < N; i++)
                        for (int k = 0; k < 3; k++) { p[i][k] *= beta; p[i][k] += r[i][k]; }
                }
                applyA(q, p, mF, bF, kF);
                const Real den = vdot(p, q);
                lastDen = den;
                if (den != 0.0) {
                    if (std::fabs(den) <= cgThreshold) {
                        if (nb_iter == 1 && timeStepCount == 0) {
                            // warning only
                        } else {
                            if (nb_iter == 1 && !equilibriumReached) equilibriumReached = true;
                            if (nb_iter > 1) equilibriumReached = false;
                            endcond = 2;
                            break;
                        }
                    }
                    alpha = rho / den;
                    // cgstep_alpha: vMultiOp -> vOp(x, x, p, alpha) and vOp(r, r, q, -alpha)
                    for (size_t i = 0; i < N; i++)
                        for (int k = 0; k < 3; k++) xs[i][k] += p[i][k] * alpha;
                    for (size_t i = 0; i < N; i++)
                        for (int k = 0; k < 3; k++) r[i][k] += q[i][k] * (-alpha);
                } else {
                    endcond = 3;
                    break;
                }
                rho_1 = rho;
            }
            if (nb_iter > cgIterations) nb_iter = cgIterations;
        } else {
            endcond = 4;
        }
        timeStepCount++;
        lastIterations = (int)nb_iter;
        lastEndCond = endcond;
    }

    // EulerImplicitSolver::solve (EulerImplicitSolver.cpp:83-300) with solveConstraint = false;
    // then DefaultAnimationLoop.
    void step(Real dt) {
        const size_t N = n();
        const Real h = dt;
        const Real tr = trapezoidal ? 0.5 : 1.0;   // :118-123
        // computeForce(f, clear, accumulate): resetForce then addForce of every force field
        f.assign(N, Vec3());
        for (ForceField* ff : forceFields()) ff->addForce(f, x, v);
        // b = f; unless firstOrder: b += (-rm M + (h tr + rs) K) v (addMBKv, B = 0), b *= h;
        // then project (:136-162)
        std::vector<Vec3> b = f;
        if (!firstOrder) {
            const Real mF = -rayleighMass, bF = 0, kF = h * tr + rayleighStiffness;
            for (ForceField* ff : forceFields()) ff->addMBKdx(b, v, mF, bF, kF);
            for (auto& e : b)
                for (int k = 0; k < 3; k++) e[k] *= h;
        }
        projectResponse(b);
        // system factors (:167-173)
        const Real mFact = firstOrder ? 1 : 1 + tr * h * rayleighMass;
        const Real bFact = firstOrder ? 0 : -tr * h;
        const Real kFact = firstOrder ? -h * tr : -tr * h * (tr * h + rayleighStiffness);
        std::vector<Vec3> xs(N);
        if (precond) {
            if (xsol.size() != N) xsol.assign(N, Vec3());
            pcgSolve(xsol, b, mFact, bFact, kFact);
            xs = xsol;
        } else if (warmStart) {
            if (xsol.size() != N) xsol.assign(N, Vec3());
            cgSolve(xsol, b, mFact, bFact, kFact, true);
            xs = xsol;
        } else {
            cgSolve(xs, b, mFact, bFact, kFact);
        }
        // the vMultiOp update (:246-268; MechanicalObject.inl:2209-2230): v = v + x (firstOrder:
        // v = x), pos = pos + v h
        const Real fxv = (Real)h;
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) {
                v[i][k] = firstOrder ? xs[i][k] : v[i][k] + xs[i][k];
                x[i][k] += v[i][k] * fxv;
            }
        // vdamping: v *= exp(-h vdamping) after the position update (:298-299)
        if (velocityDamping != 0.0) {
            const Real d = std::exp(-h * velocityDamping);
            for (size_t i = 0; i < N; i++)
                for (int k = 0; k < 3; k++) v[i][k] *= d;
        }
        // projectPositionAndVelocity (DefaultAnimationLoop.cpp:228-234) with the context time
        // still at the step start, the fixed constraint before the oscillator: the fixed nodes'
        // velocities cleared if activate_projectVelocity (FixedProjectiveConstraint.inl:237-258;
        // its projectPosition is a no-op), then OscillatorProjectiveConstraint.inl:77-106
        if (projectVelocity) {
            if (fixAll) {
                for (auto& e : v) e = Vec3();
            } else {
                for (int i : fixed) v[i] = Vec3();
            }
        }
        for (const auto& o : oscillators) {
            const Real t = time;
            for (int k = 0; k < 3; k++) {
                x[o.index][k] = o.mean[k] + o.amplitude[k] * std::sin(o.pulsation * t + o.phase);
                v[o.index][k] = o.amplitude[k] * o.pulsation * std::cos(o.pulsation * t + o.phase);
            }
        }
        // LinearMovementProjectiveConstraint: projectPosition, then projectVelocity
        // (MechanicalProjectPositionAndVelocityVisitor.cpp:35-40)
        for (auto& lm : movements) {
            lm.projectPosition(x, time);
            lm.projectVelocity(v, time);
        }
        time += dt;
    }
};

} // namespace oracle

// ==========================================================================================
// C interface for the tests (ctypes)
// ==========================================================================================
using namespace oracle;

extern "C" {

void* orc_create(int n_nodes, const double* positions) {
    Scene* s = new Scene();
    s->x.resize(n_nodes);
    for (int i = 0; i < n_nodes; i++) s->x[i] = Vec3(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]);
    s->x0 = s->x;
    s->v.assign(n_nodes, Vec3());
    s->f.assign(n_nodes, Vec3());
    return s;
}

void orc_destroy(void* h) { delete static_cast<Scene*>(h); }

void orc_set_gravity(void* h, double gx, double gy, double gz) {
    Scene* s = static_cast<Scene*>(h);
    s->gravity = Vec3(gx, gy, gz);
    if (s->mass) s->mass->gravity = s->gravity;
}

int orc_set_fem(void* h, int n_tets, const int* tets, const double* young, int n_young,
                const double* poisson, int n_poisson, int method) {
    Scene* s = static_cast<Scene*>(h);
    std::vector<Tetra> t(n_tets);
    for (int i = 0; i < n_tets; i++)
        for (int k = 0; k < 4; k++) {
            t[i][k] = tets[4 * i + k];
            if (t[i][k] < 0 || (size_t)t[i][k] >= s->n()) return -1;
        }
    s->fem.reset(new TetraFEM());
    s->fem->init(s->x0, t, std::vector<Real>(young, young + n_young), std::vector<Real>(poisson, poisson + n_poisson), method);
    return 0;
}

void orc_set_uniform_mass(void* h, double vertex_mass) {
    Scene* s = static_cast<Scene*>(h);
    UniformMass* m = new UniformMass();
    m->vertexMass = vertex_mass;
    m->gravity = s->gravity;
    s->mass.reset(m);
}

// UniformMass::initFromTotalMass (UniformMass.inl:329-345): vertexMass = totalMass / Real(n)
void orc_set_uniform_total_mass(void* h, double total_mass) {
    Scene* s = static_cast<Scene*>(h);
    orc_set_uniform_mass(h, total_mass / Real(s->n()));
}

void orc_set_diagonal_mass(void* h, double density) {
    Scene* s = static_cast<Scene*>(h);
    DiagonalMass* m = new DiagonalMass();
    m->init(s->x0, s->fem ? s->fem->elems : std::vector<Tetra>(), density);
    m->gravity = s->gravity;
    s->mass.reset(m);
}

void orc_add_plane(void* h, const double* normal, double d, double stiffness, double damping, double max_force,
                   int bilateral) {
    Scene* s = static_cast<Scene*>(h);
    PlaneForceField* pf = new PlaneForceField();
    pf->normal = Vec3(normal[0], normal[1], normal[2]);
    pf->d = d; pf->stiffness = stiffness; pf->damping = damping; pf->maxForce = max_force;
    pf->bilateral = bilateral != 0;
    s->extra.emplace_back(pf);
}

void orc_set_diagonal_total_mass(void* h, double total_mass) {
    Scene* s = static_cast<Scene*>(h);
    DiagonalMass* m = new DiagonalMass();
    m->init(s->x0, s->fem ? s->fem->elems : std::vector<Tetra>(), total_mass, true);
    m->gravity = s->gravity;
    s->mass.reset(m);
}

void orc_add_axis_spring(void* h, int node, int axis, double stiffness, double anchor) {
    Scene* s = static_cast<Scene*>(h);
    AxisSpring* sp = new AxisSpring();
    sp->node = node; sp->axis = axis; sp->stiffness = stiffness; sp->anchor = anchor;
    s->extra.emplace_back(sp);
}

void orc_set_fixed(void* h, const int* idx, int n, int fix_all) {
    Scene* s = static_cast<Scene*>(h);
    s->fixed.assign(idx, idx + n);
    s->fixAll = fix_all != 0;
}

void orc_add_linear_movement(void* h, const int* idx, int n, const double* key_times, const double* moves, int nk,
                             int relative) {
    Scene* s = static_cast<Scene*>(h);
    LinearMovement lm;
    lm.indices.assign(idx, idx + n);
    lm.keyTimes.assign(key_times, key_times + nk);
    for (int k = 0; k < nk; k++) lm.movements.push_back(Vec3(moves[3 * k], moves[3 * k + 1], moves[3 * k + 2]));
    lm.relative = relative != 0;
    s->movements.push_back(lm);
}
void orc_set_preconditioner(void* h, int kind) { static_cast<Scene*>(h)->precond = kind; }
void orc_set_warm_start(void* h, int on) { static_cast<Scene*>(h)->warmStart = on != 0; }
// the preconditioner's inverted 3x3 blocks (Jacobi: the inverted diagonals) for factors mF, kF,
// with the current rotations: out 9 per node, row-major
void orc_precond_blocks(void* h, double mF, double kF, double* out) {
    Scene* s = static_cast<Scene*>(h);
    std::vector<Mat3> P;
    s->precondInverse(P, mF, kF);
    for (size_t i = 0; i < P.size(); i++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) out[9 * i + 3 * a + b] = P[i].m[a][b];
}
void orc_set_project_velocity(void* h, int on) { static_cast<Scene*>(h)->projectVelocity = on != 0; }

void orc_add_oscillator(void* h, int index, const double* mean, const double* amplitude, double pulsation, double phase) {
    Scene* s = static_cast<Scene*>(h);
    Oscillator o;
    o.index = index;
    o.mean = Vec3(mean[0], mean[1], mean[2]);
    o.amplitude = Vec3(amplitude[0], amplitude[1], amplitude[2]);
    o.pulsation = pulsation;
    o.phase = phase;
    s->oscillators.push_back(o);
}

void orc_set_solver(void* h, double rayleigh_stiffness, double rayleigh_mass, int cg_iterations,
                    double cg_tolerance, double cg_threshold) {
    Scene* s = static_cast<Scene*>(h);
    s->rayleighStiffness = rayleigh_stiffness;
    s->rayleighMass = rayleigh_mass;
    s->cgIterations = (unsigned)cg_iterations;
    s->cgTolerance = cg_tolerance < 0 ? 1e-5 : cg_tolerance;     // CGLinearSolver::init :49-69
    s->cgThreshold = cg_threshold < 0 ? 1e-5 : cg_threshold;
}

void orc_set_integrator(void* h, int first_order, int trapezoidal, double vdamping) {
    Scene* s = static_cast<Scene*>(h);
    s->firstOrder = first_order != 0;
    s->trapezoidal = trapezoidal != 0;
    s->velocityDamping = vdamping;
}

void orc_set_dot_mode(void* h, int mode) { static_cast<Scene*>(h)->dotMode = mode; }

void orc_set_threads(void* h, int nthreads) {
    Scene* s = static_cast<Scene*>(h);
    if (s->fem) s->fem->nthreads = nthreads < 1 ? 1 : nthreads;
}

void orc_step(void* h, double dt, int nsteps) {
    Scene* s = static_cast<Scene*>(h);
    for (int i = 0; i < nsteps; i++) s->step(dt);
}

// which: 0 position, 1 velocity, 2 force (last computeForce), 3 rest position
void orc_get_vec(void* h, int which, double* out) {
    Scene* s = static_cast<Scene*>(h);
    const std::vector<Vec3>* src = which == 0 ? &s->x : which == 1 ? &s->v : which == 2 ? &s->f : &s->x0;
    for (size_t i = 0; i < src->size(); i++)
        for (int k = 0; k < 3; k++) out[3 * i + k] = (*src)[i][k];
}

void orc_set_vec(void* h, int which, const double* in) {
    Scene* s = static_cast<Scene*>(h);
    std::vector<Vec3>* dst = which == 0 ? &s->x : &s->v;
    for (size_t i = 0; i < dst->size(); i++)
        for (int k = 0; k < 3; k++) (*dst)[i][k] = in[3 * i + k];
}

// stats: [iterations, endcond, timeStepCount], [lastErr, lastDen, time]
void orc_get_stats(void* h, int* istats, double* dstats) {
    Scene* s = static_cast<Scene*>(h);
    istats[0] = s->lastIterations;
    istats[1] = s->lastEndCond;
    istats[2] = s->timeStepCount;
    dstats[0] = s->lastErr;
    dstats[1] = s->lastDen;
    dstats[2] = s->time;
}

// Component-level calls, the per-call unit of TetrahedronFEMForceField:
// f += addForce(x) (updates the rotations); df += addDForce(dx, kFactor)
void orc_fem_add_force(void* h, const double* xin, double* fio) {
    Scene* s = static_cast<Scene*>(h);
    const size_t N = s->n();
    std::vector<Vec3> xx(N), ff(N), vv(N);
    for (size_t i = 0; i < N; i++) {
        xx[i] = Vec3(xin[3 * i], xin[3 * i + 1], xin[3 * i + 2]);
        ff[i] = Vec3(fio[3 * i], fio[3 * i + 1], fio[3 * i + 2]);
    }
    s->fem->addForce(ff, xx, vv);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) fio[3 * i + k] = ff[i][k];
}

void orc_fem_add_dforce(void* h, const double* dxin, double* dfio, double kFactor) {
    Scene* s = static_cast<Scene*>(h);
    const size_t N = s->n();
    std::vector<Vec3> dx(N), df(N);
    for (size_t i = 0; i < N; i++) {
        dx[i] = Vec3(dxin[3 * i], dxin[3 * i + 1], dxin[3 * i + 2]);
        df[i] = Vec3(dfio[3 * i], dfio[3 * i + 1], dfio[3 * i + 2]);
    }
    s->fem->addDForce(df, dx, (Real)kFactor);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) dfio[3 * i + k] = df[i][k];
}

// element data as the reference stores it: R0 (_initialRotations), R (rotations),
// K (6x6), J (12x6), rotated rest coordinates (4x3)
void orc_fem_get_element(void* h, int e, double* R0, double* R, double* K, double* J, double* rest) {
    Scene* s = static_cast<Scene*>(h);
    TetraFEM* fem = s->fem.get();
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            if (fem->method == TetraFEM::LARGE) {
                R0[3 * i + j] = fem->initialRotations[e].m[i][j];
                R[3 * i + j] = fem->rotations[e].m[i][j];
            } else {
                R0[3 * i + j] = R[3 * i + j] = (i == j) ? 1.0 : 0.0;
            }
        }
    for (int i = 0; i < 36; i++) K[i] = fem->K[e][i];
    for (int i = 0; i < 72; i++) J[i] = fem->J[e][i];
    for (int k = 0; k < 4; k++)
        for (int q = 0; q < 3; q++)
            rest[3 * k + q] = (fem->method == TetraFEM::LARGE) ? fem->rotatedInitial[e][k][q] : fem->initialPoints[fem->elems[e][k]][q];
}

double orc_fem_rest_volume(void* h) { return static_cast<Scene*>(h)->fem->restVolume; }

// addKToMatrix with the current rotations (those of the last addForce or step) into a dense
// row-major 3N x 3N matrix that the caller zeroed
void orc_stiffness_dense(void* h, double k, double* mat) {
    Scene* s = static_cast<Scene*>(h);
    s->fem->addKToMatrix(mat, s->n(), (Real)k);
}

// computeVonMisesStress from the current positions (mode 1 corotational, 2 Green strain):
// per element (n_tets) and per node (n_nodes); returns 0, or -1 when a rest element's matrix
// has no inverse or the mode / method does not allow it
int orc_von_mises(void* h, int mode, double* per_elem, double* per_node) {
    Scene* s = static_cast<Scene*>(h);
    TetraFEM* fem = s->fem.get();
    if (!fem || (mode != 1 && mode != 2) || (mode == 1 && fem->method != TetraFEM::LARGE)) return -1;
    if (fem->elemShapeFun.size() != fem->elems.size() && !fem->initVonMises()) return -1;
    std::vector<Real> vme, vmn;
    fem->computeVonMisesStress(mode, s->x, vme, vmn);
    for (size_t e = 0; e < vme.size(); e++) per_elem[e] = vme[e];
    for (size_t n = 0; n < vmn.size(); n++) per_node[n] = vmn[n];
    return 0;
}

void orc_get_masses(void* h, double* out) {
    Scene* s = static_cast<Scene*>(h);
    if (auto* d = dynamic_cast<DiagonalMass*>(s->mass.get())) {
        for (size_t i = 0; i < d->masses.size(); i++) out[i] = d->masses[i];
    } else if (auto* u = dynamic_cast<UniformMass*>(s->mass.get())) {
        for (size_t i = 0; i < s->n(); i++) out[i] = u->vertexMass;
    } else if (auto* mm = dynamic_cast<MeshMatrixMass*>(s->mass.get())) {
        for (size_t i = 0; i < mm->vertexMass.size(); i++) out[i] = mm->vertexMass[i];
    }
}

void orc_set_meshmatrix_mass(void* h, const double* density, int n_density, int lumping) {
    Scene* s = static_cast<Scene*>(h);
    MeshMatrixMass* m = new MeshMatrixMass();
    m->init(s->x0, s->fem ? s->fem->elems : std::vector<Tetra>(), std::vector<Real>(density, density + n_density),
            lumping != 0);
    m->gravity = s->gravity;
    s->mass.reset(m);
}

// component-level calls on caller vectors (n x 3): Mass::addMDx / accFromF / addForce, the
// projective constraints, vDot, CGLinearSolver::solve
static std::vector<Vec3> to_vec(const double* a, size_t n) {
    std::vector<Vec3> r(n);
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) r[i][k] = a[3 * i + k];
    return r;
}
static void from_vec(const std::vector<Vec3>& r, double* a) {
    for (size_t i = 0; i < r.size(); i++)
        for (int k = 0; k < 3; k++) a[3 * i + k] = r[i][k];
}

void orc_mass_add_mdx(void* h, const double* dx, double* res, double factor) {
    Scene* s = static_cast<Scene*>(h);
    std::vector<Vec3> r = to_vec(res, s->n());
    s->mass->addMDx(r, to_vec(dx, s->n()), factor);
    from_vec(r, res);
}

int orc_mass_acc_from_f(void* h, const double* f, double* a) {
    Scene* s = static_cast<Scene*>(h);
    std::vector<Vec3> r(s->n());
    if (!s->mass->accFromF(r, to_vec(f, s->n()))) return -1;
    from_vec(r, a);
    return 0;
}

void orc_mass_add_force(void* h, double* f) {
    Scene* s = static_cast<Scene*>(h);
    std::vector<Vec3> r = to_vec(f, s->n());
    s->mass->addForce(r, s->x, s->v);
    from_vec(r, f);
}

// what: 0 projectResponse, 1 projectVelocity at t, 2 projectPosition at t
void orc_project(void* h, int what, double* vec, double t) {
    Scene* s = static_cast<Scene*>(h);
    std::vector<Vec3> r = to_vec(vec, s->n());
    if (what == 0) s->projectResponse(r);
    else if (what == 1) s->projectVelocityAt(r, t);
    else s->projectPositionAt(r, t);
    from_vec(r, vec);
}

// mode 0: the reference's sequential order; 1: the GPU's blocked order
double orc_vdot(void* h, const double* a, const double* b, int mode) {
    Scene* s = static_cast<Scene*>(h);
    const std::vector<Vec3> va = to_vec(a, s->n()), vb = to_vec(b, s->n());
    return mode == 1 ? s->dot_blocked(va, vb) : s->dot_ref(va, vb);
}

// CGLinearSolver::solve of (mF M + bF B + kF K) x = b with the scene's CG Data; the rotations of
// the last addForce (orc_fem_add_force or a step)
void orc_cg_solve(void* h, const double* b, double* x, double mF, double bF, double kF) {
    Scene* s = static_cast<Scene*>(h);
    std::vector<Vec3> xs(s->n());
    s->cgSolve(xs, to_vec(b, s->n()), mF, bF, kF);
    from_vec(xs, x);
}

int orc_get_edge_count(void* h) {
    Scene* s = static_cast<Scene*>(h);
    auto* m = dynamic_cast<MeshMatrixMass*>(s->mass.get());
    return m ? (int)m->edges.size() : 0;
}

void orc_get_edges(void* h, int* edges, double* edge_mass) {
    Scene* s = static_cast<Scene*>(h);
    auto* m = dynamic_cast<MeshMatrixMass*>(s->mass.get());
    if (!m) return;
    for (size_t j = 0; j < m->edges.size(); j++) {
        edges[2 * j] = m->edges[j][0];
        edges[2 * j + 1] = m->edges[j][1];
        edge_mass[j] = m->edgeMass[j];
    }
}

} // extern "C"
