// Host mesh builder.  Connectivity uses hash maps (O(E)) but reproduces the
// numbering of the reference's O(E^2) linked-list build exactly:
//   * edges are numbered in first-appearance order over (element, face 1..4)
//     with the node pair of the element that created them (MeshFile.f90:287-362);
//   * the LEFT element of an edge is the lowest element id that owns it
//     (MeshFile.f90:499-574,754-783);
//   * periodic zones are paired by coordinate match, first zone first
//     (QuadMesh_Zones.incf:22-131).
// Geometry follows the transfinite mapping of QuadElement_QuadMappings.incf:22-367
// and the edge curves of QuadElement_EdgeMappings.incf:22-442.
#include "hg_mesh.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>

namespace hg {

namespace {

// Utilities.f90:37-64 (Kopriva's AlmostEqual)
bool almost_equal(double a, double b) {
    if (a == 0.0 || b == 0.0) return std::fabs(a - b) <= 2.0 * DBL_EPSILON;
    return std::fabs(b - a) <= 2.0 * DBL_EPSILON * std::fmax(std::fabs(a), std::fabs(b));
}

// Legendre polynomial L_n and derivative by the three-term recursion (Kopriva, Alg. 22).
void legendre(int n, double x, double &L, double &dL) {
    if (n == 0) { L = 1.0; dL = 0.0; return; }
    if (n == 1) { L = x; dL = 1.0; return; }
    double l2 = 1.0, d2 = 0.0, l1 = x, d1 = 1.0;
    for (int k = 2; k <= n; ++k) {
        L = ((2.0 * k - 1.0) * x * l1 - (k - 1.0) * l2) / k;
        dL = d2 + (2.0 * k - 1.0) * l1;
        l2 = l1; l1 = L; d2 = d1; d1 = dL;
    }
}

// Gauss-Legendre nodes/weights on [-1,1] by Newton iteration (Kopriva, Alg. 23).
void gauss_nodes(int N, double *x, double *w) {
    constexpr double kPi = 3.141592653589793238462643;
    if (N == 0) { x[0] = 0.0; w[0] = 2.0; return; }
    if (N == 1) {
        x[0] = -std::sqrt(1.0 / 3.0); w[0] = 1.0; x[1] = -x[0]; w[1] = 1.0;
    } else {
        const double tol = 4.0 * DBL_EPSILON;
        for (int j = 0; j < (N + 1) / 2; ++j) {
            double xj = -std::cos((2 * j + 1) * kPi / (2 * N + 2)), L, dL;
            for (int it = 0; it <= 100; ++it) {
                legendre(N + 1, xj, L, dL);
                double d = -L / dL;
                xj += d;
                if (std::fabs(d) <= tol * std::fabs(xj)) break;
            }
            legendre(N + 1, xj, L, dL);
            x[j] = xj; x[N - j] = -xj;
            w[j] = 2.0 / ((1.0 - xj * xj) * dL * dL); w[N - j] = w[j];
        }
    }
    if (N % 2 == 0) {
        double L, dL;
        legendre(N + 1, 0.0, L, dL);
        x[N / 2] = 0.0; w[N / 2] = 2.0 / (dL * dL);
    }
}

void bary_weights(int N, const double *x, double *wb) {
    for (int j = 0; j <= N; ++j) wb[j] = 1.0;
    for (int j = 1; j <= N; ++j)
        for (int k = 0; k < j; ++k) { wb[k] *= (x[k] - x[j]); wb[j] *= (x[j] - x[k]); }
    for (int j = 0; j <= N; ++j) wb[j] = 1.0 / wb[j];
}

void lagrange_bary(double x, int N, const double *nodes, const double *wb, double *p) {
    bool hit = false;
    for (int j = 0; j <= N; ++j) {
        p[j] = 0.0;
        if (almost_equal(x, nodes[j])) { p[j] = 1.0; hit = true; }
    }
    if (hit) return;
    double d = 0.0;
    for (int j = 0; j <= N; ++j) { p[j] = wb[j] / (x - nodes[j]); d += p[j]; }
    for (int j = 0; j <= N; ++j) p[j] /= d;
}

// T[k][j] maps values at old nodes (N+1) to new nodes (M+1) (Kopriva, Alg. 32).
void interpolation_matrix(int N, int M, const double *xo, const double *wb, const double *xn, double *T) {
    for (int k = 0; k <= M; ++k) {
        bool hit = false;
        for (int j = 0; j <= N; ++j) {
            T[k * (N + 1) + j] = 0.0;
            if (almost_equal(xn[k], xo[j])) { T[k * (N + 1) + j] = 1.0; hit = true; }
        }
        if (hit) continue;
        double s = 0.0;
        for (int j = 0; j <= N; ++j) { double t = wb[j] / (xn[k] - xo[j]); T[k * (N + 1) + j] = t; s += t; }
        for (int j = 0; j <= N; ++j) T[k * (N + 1) + j] /= s;
    }
}

void make_spectral(Spectral &s, int N) {
    s.N = N; s.P = N + 1;
    gauss_nodes(N, s.xi, s.w);
    for (int i = 0; i <= N; ++i) { s.xi[i] = 0.5 + 0.5 * s.xi[i]; s.w[i] = 0.5 * s.w[i]; }
    bary_weights(N, s.xi, s.wb);
    double bw[kMaxP];
    bary_weights(N, s.xi, bw);
    for (int i = 0; i <= N; ++i) {
        s.D[i][i] = 0.0;
        for (int j = 0; j <= N; ++j)
            if (j != i) { s.D[i][j] = bw[j] / (bw[i] * (s.xi[i] - s.xi[j])); s.D[i][i] -= s.D[i][j]; }
    }
    for (int i = 0; i <= N; ++i)
        for (int l = 0; l <= N; ++l) s.hatD[i][l] = (1.0 / s.w[i]) * (s.w[l] * s.D[l][i]);
    lagrange_bary(0.0, N, s.xi, s.wb, s.lb[0]);
    lagrange_bary(1.0, N, s.xi, s.wb, s.lb[1]);
    for (int side = 0; side < 2; ++side)
        for (int i = 0; i <= N; ++i) {
            s.lbw[side][i] = s.lb[side][i] / s.w[i];
            double acc = 0.0;
            for (int j = 0; j <= N; ++j) acc += s.D[j][i] * s.lb[side][j];
            s.dlb[side][i] = acc;
            s.dlbw[side][i] = acc / s.w[i];
        }
}

struct PairHash {
    size_t operator()(uint64_t k) const { return std::hash<uint64_t>()(k * 0x9E3779B97F4A7C15ull); }
};
inline uint64_t key(int a, int b) {
    uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    return ((uint64_t)hi << 32) | lo;
}

// searchEdge / compareEdges (QuadElement_Auxiliars.incf:27-117): position 1..4,
// direction +1/-1 relative to the element's local parameter direction.
void locate(const int *en, const int *edn, int &pos, int &dir) {
    for (int p = 1; p <= 4; ++p) {
        int c0 = en[p - 1], c1 = en[p % 4];
        bool fwd_pos = (p == 1 || p == 2);
        if (c0 == edn[0] && c1 == edn[1]) { pos = p; dir = fwd_pos ? 1 : -1; return; }
        if (c0 == edn[1] && c1 == edn[0]) { pos = p; dir = fwd_pos ? -1 : 1; return; }
    }
    pos = -1; dir = 0;
}

// Edge_Invert (QuadElement_EdgeProcedures.incf:433-521)
void invert(Edge &e, bool flip_normal) {
    int P = e.N + 1;
    std::swap(e.nodes[0], e.nodes[1]);
    for (int i = 0; i < P / 2; ++i)
        for (int d = 0; d < 2; ++d) std::swap(e.X[2 * i + d], e.X[2 * (P - 1 - i) + d]);
    double s = flip_normal ? -1.0 : 1.0;
    if (!e.curved) {
        e.dX[0] = -e.dX[0]; e.dX[1] = -e.dX[1];
        e.n[0] *= s; e.n[1] *= s;
        return;
    }
    std::vector<double> dX = e.dX, n = e.n, dS = e.dS;
    for (int i = 0; i < P; ++i) {
        int r = P - 1 - i;
        e.dX[2 * i] = -dX[2 * r]; e.dX[2 * i + 1] = -dX[2 * r + 1];
        e.dS[i] = dS[r];
        e.n[2 * i] = s * n[2 * r]; e.n[2 * i + 1] = s * n[2 * r + 1];
    }
}

// Edge_SetCurve / CurvedEdge_SetCurve (QuadElement_EdgeMappings.incf:22-50,185-260)
void set_curve(const Mesh &m, Edge &e, const double *cpts, int order) {
    const Spectral &s = spectral(e.N);
    int P = e.N + 1;
    e.X.assign(2 * P, 0.0); e.dX.assign(2 * P, 0.0); e.n.assign(2 * P, 0.0); e.dS.assign(P, 0.0);
    const double *a = &m.pts[2 * e.nodes[0]], *b = &m.pts[2 * e.nodes[1]];
    if (!e.curved) {
        for (int p = 0; p < P; ++p)
            for (int d = 0; d < 2; ++d) e.X[2 * p + d] = a[d] * (1.0 - s.xi[p]) + b[d] * s.xi[p];
        double sx = b[1] - a[1], sy = a[0] - b[0];
        e.dS[0] = std::sqrt(sx * sx + sy * sy);
        e.n[0] = sx / e.dS[0]; e.n[1] = sy / e.dS[0];
        e.dX[0] = b[0] - a[0]; e.dX[1] = b[1] - a[1];
        return;
    }
    constexpr double kPi = 3.141592653589793238462643;
    int NC = e.NCurve;
    std::vector<double> cgl(order + 1), cgi(NC + 1), wb(kMaxP + 4), T((kMaxP + 4) * (kMaxP + 4)), Xi(2 * (NC + 1));
    for (int k = 0; k <= order; ++k) cgl[k] = 0.5 + 0.5 * std::cos(kPi * (order - k) / (1.0 * order));
    for (int k = 0; k <= NC; ++k) cgi[k] = 0.5 + 0.5 * std::cos(kPi * (NC - k) / (1.0 * NC));
    bary_weights(order, cgl.data(), wb.data());
    interpolation_matrix(order, NC, cgl.data(), wb.data(), cgi.data(), T.data());
    for (int k = 0; k <= NC; ++k)
        for (int d = 0; d < 2; ++d) {
            double acc = 0.0;
            for (int j = 0; j <= order; ++j) acc += cpts[2 * j + d] * T[k * (order + 1) + j];
            Xi[2 * k + d] = acc;
        }
    bary_weights(NC, cgi.data(), wb.data());
    interpolation_matrix(NC, e.N, cgi.data(), wb.data(), s.xi, T.data());
    for (int i = 0; i < P; ++i)
        for (int d = 0; d < 2; ++d) {
            double acc = 0.0;
            for (int k = 0; k <= NC; ++k) acc += Xi[2 * k + d] * T[i * (NC + 1) + k];
            e.X[2 * i + d] = acc;
        }
    for (int i = 0; i < P; ++i)
        for (int d = 0; d < 2; ++d) {
            double acc = 0.0;
            for (int j = 0; j < P; ++j) acc += e.X[2 * j + d] * s.D[i][j];
            e.dX[2 * i + d] = acc;
        }
    for (int i = 0; i < P; ++i) {
        e.dS[i] = std::sqrt(e.dX[2 * i] * e.dX[2 * i] + e.dX[2 * i + 1] * e.dX[2 * i + 1]);
        e.n[2 * i] = e.dX[2 * i + 1] / e.dS[i];
        e.n[2 * i + 1] = -e.dX[2 * i] / e.dS[i];
    }
}

// EvaluateX forwards at x: Edge_AnalyticalX from the nodes, CurvedEdge_InterpolantX from the
// edge's points (QuadElement_EdgeMappings.incf:52-78,355-380).
void point_at(const Mesh &m, const Edge &e, double x, double *X) {
    if (!e.curved) {
        const double *a = &m.pts[2 * e.nodes[0]], *b = &m.pts[2 * e.nodes[1]];
        for (int d = 0; d < 2; ++d) X[d] = a[d] * (1.0 - x) + b[d] * x;
        return;
    }
    const Spectral &s = spectral(e.N);
    double lj[kMaxP];
    lagrange_bary(x, e.N, s.xi, s.wb, lj);
    for (int d = 0; d < 2; ++d) {
        double p = 0.0;
        for (int j = 0; j <= e.N; ++j) p += e.X[2 * j + d] * lj[j];
        X[d] = p;
    }
}

// Edge point and tangent at the local index of an element of order elN
// (QuadElement_EdgeMappings.incf:52-177,262-442): an interior edge is evaluated at the element's
// node (QuadElement_QuadMappings.incf:55-62), which differs from the edge's nodes when the edge
// takes the higher order of its two elements.  A subdivided edge is evaluated forwards, on
// [0.5, 1] for its NORTH element (qpos 1) and [0, 0.5] for its SOUTH one (qpos 2), at the
// element's nodes reversed when it runs its half backwards, dX scaled by the half and the
// direction (QuadElement_QuadMappings.incf:66-115).
void edge_point(const Mesh &m, const Edge &e, int elN, int idx, int dir, int qpos, double *X, double *dX) {
    const Spectral &s = spectral(e.N);
    if (e.hang >= 0) {
        const double *xs = spectral(elN).xi;
        double xv = xs[idx], scal = 1.0;
        if (qpos == 1) { xv = 0.5 + 0.5 * (dir == -1 ? xs[elN - idx] : xs[idx]); scal = dir == -1 ? -0.5 : 0.5; }
        else if (qpos == 2) { xv = 0.5 * (dir == -1 ? xs[elN - idx] : xs[idx]); scal = dir == -1 ? -0.5 : 0.5; }
        point_at(m, e, xv, X);
        if (!e.curved) {
            const double *a = &m.pts[2 * e.nodes[0]], *b = &m.pts[2 * e.nodes[1]];
            for (int d = 0; d < 2; ++d) dX[d] = scal * (b[d] - a[d]);
        } else {
            double lj[kMaxP];
            lagrange_bary(xv, e.N, s.xi, s.wb, lj);
            for (int d = 0; d < 2; ++d) {
                double q = 0.0;
                for (int j = 0; j <= e.N; ++j) q += e.dX[2 * j + d] * lj[j];
                dX[d] = scal * q;
            }
        }
        return;
    }
    if (e.bdry) {
        int c = dir == 1 ? idx : e.N - idx;
        X[0] = e.X[2 * c]; X[1] = e.X[2 * c + 1];
        int cd = e.curved ? c : 0;
        dX[0] = (double)dir * e.dX[2 * cd]; dX[1] = (double)dir * e.dX[2 * cd + 1];
        return;
    }
    const double xe = spectral(elN).xi[idx];
    double x = dir == -1 ? 1.0 - xe : xe;
    if (!e.curved) {
        const double *a = &m.pts[2 * e.nodes[0]], *b = &m.pts[2 * e.nodes[1]];
        for (int d = 0; d < 2; ++d) { X[d] = a[d] * (1.0 - x) + b[d] * x; dX[d] = (double)dir * (b[d] - a[d]); }
        return;
    }
    double lj[kMaxP];
    lagrange_bary(x, e.N, s.xi, s.wb, lj);
    for (int d = 0; d < 2; ++d) {
        double p = 0.0, q = 0.0;
        for (int j = 0; j <= e.N; ++j) { p += e.X[2 * j + d] * lj[j]; q += e.dX[2 * j + d] * lj[j]; }
        X[d] = p; dX[d] = (double)dir * q;
    }
}

void element_mappings(const Mesh &m, Elem &el) {
    const Spectral &s = spectral(el.N);
    int P = el.P;
    std::vector<double> XE(2 * P * P * 4), DXE(2 * P * P * 4);
    auto at = [P](int d, int i, int j, int f) { return ((f * P + j) * P + i) * 2 + d; };
    for (int pos = 1; pos <= 4; ++pos) {
        const Edge &e = m.ed[el.edges[pos - 1]];
        bool horiz = (pos == 1 || pos == 3);
        for (int a = 0; a < P; ++a) {
            double X[2], dX[2];
            edge_point(m, e, el.N, a, el.dir[pos - 1], el.qpos[pos - 1], X, dX);
            for (int b = 0; b < P; ++b) {
                int i = horiz ? a : b, j = horiz ? b : a;
                for (int d = 0; d < 2; ++d) { XE[at(d, i, j, pos - 1)] = X[d]; DXE[at(d, i, j, pos - 1)] = dX[d]; }
            }
        }
    }
    const double *n1 = &m.pts[2 * el.nodes[0]], *n2 = &m.pts[2 * el.nodes[1]];
    const double *n3 = &m.pts[2 * el.nodes[2]], *n4 = &m.pts[2 * el.nodes[3]];
    el.X.assign(2 * P * P, 0.0); el.Ja.assign(4 * P * P, 0.0); el.jac.assign(P * P, 0.0);
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < P; ++j) {
            double xi = s.xi[i], eta = s.xi[j], dx[2][2];
            for (int d = 0; d < 2; ++d) {
                el.X[(i + P * j) * 2 + d] = (1.0 - eta) * XE[at(d, i, j, 0)] + xi * XE[at(d, i, j, 1)] +
                                            eta * XE[at(d, i, j, 2)] + (1.0 - xi) * XE[at(d, i, j, 3)] -
                                            n1[d] * (1.0 - xi) * (1.0 - eta) - n2[d] * xi * (1.0 - eta) -
                                            n3[d] * xi * eta - n4[d] * (1.0 - xi) * eta;
                dx[d][0] = (1.0 - eta) * DXE[at(d, i, j, 0)] + XE[at(d, i, j, 1)] + eta * DXE[at(d, i, j, 2)] -
                           XE[at(d, i, j, 3)] + n1[d] * (1.0 - eta) - n2[d] * (1.0 - eta) - n3[d] * eta + n4[d] * eta;
                dx[d][1] = -XE[at(d, i, j, 0)] + xi * DXE[at(d, i, j, 1)] + XE[at(d, i, j, 2)] +
                           (1.0 - xi) * DXE[at(d, i, j, 3)] + n1[d] * (1.0 - xi) + n2[d] * xi - n3[d] * xi -
                           n4[d] * (1.0 - xi);
            }
            int k = i + P * j, PP = P * P;
            el.jac[k] = dx[0][0] * dx[1][1] - dx[0][1] * dx[1][0];
            el.Ja[0 * PP + k] = dx[1][1];
            el.Ja[1 * PP + k] = -dx[1][0];
            el.Ja[2 * PP + k] = -dx[0][1];
            el.Ja[3 * PP + k] = dx[0][0];
        }
    double vol = 0.0;
    for (int j = 0; j < P; ++j) {
        double acc = 0.0;
        for (int i = 0; i < P; ++i) acc += s.w[i] * el.jac[i + P * j];
        vol += acc * s.w[j];
    }
    el.volume = vol;
}

// mergeDividedFaces (MeshFile.f90:364-490) on the edge list.  Entries are in list order (their
// ids: a removal keeps the order, a merged entry is appended), each two node ids, or three after
// a merge, and the file's boundary attribute.  For the nodes in ascending order: a node in
// exactly three entries, none of them boundary edges, two of whose neighbours are joined by an
// entry of their own (the first eight entries of each neighbour are looked at) hangs on that
// edge: the edge and its two halves are removed and [p0, p1, node] is appended.
struct EdgeList {
    std::vector<std::array<int, 3>> v;
    std::vector<char> attr, alive;
    std::vector<std::vector<int>> of_node;                          // entry ids, ascending
    std::unordered_map<uint64_t, std::vector<int>, PairHash> pair;  // two-node entries
};

void merge_divided(int nnodes, EdgeList &L) {
    auto remove = [&](int id) {
        if (id < 0) return;
        L.alive[id] = 0;
        for (int c = 0; c < 3; ++c) {
            const int n = L.v[id][c];
            if (n < 0) continue;
            auto &o = L.of_node[n];
            o.erase(std::find(o.begin(), o.end(), id));
        }
        if (L.v[id][2] < 0) {
            auto &p = L.pair[key(L.v[id][0], L.v[id][1])];
            p.erase(std::find(p.begin(), p.end(), id));
        }
    };
    auto first_pair = [&](int a, int b) {
        auto it = L.pair.find(key(a, b));
        return (it == L.pair.end() || it->second.empty()) ? -1 : it->second.front();
    };
    for (int node = 0; node < nnodes; ++node) {
        if (L.of_node[node].size() != 3) continue;
        const int inc[3] = {L.of_node[node][0], L.of_node[node][1], L.of_node[node][2]};
        if (L.attr[inc[0]] || L.attr[inc[1]] || L.attr[inc[2]]) continue;
        int pts[3];
        for (int j = 0; j < 3; ++j) {
            const auto &v = L.v[inc[j]];
            pts[j] = v[0] != node ? v[0] : (v[1] != node ? v[1] : -1);
        }
        bool done = false;
        for (int j = 0; j < 3 && !done; ++j) {
            if (pts[j] < 0) continue;
            const auto &on = L.of_node[pts[j]];
            const std::vector<int> ne(on.begin(), on.begin() + std::min<size_t>(8, on.size()));
            for (size_t k = 0; k < ne.size() && !done; ++k) {
                const int p0 = L.v[ne[k]][0], p1 = L.v[ne[k]][1];
                int point = -1;
                if (p0 != pts[j] && p0 != node) point = p0;
                else if (p1 != pts[j] && p1 != node) point = p1;
                if (point == -1 || (point != pts[0] && point != pts[1] && point != pts[2])) continue;
                remove(first_pair(p0, p1));
                remove(first_pair(p0, node));
                remove(first_pair(p1, node));
                const int id = (int)L.v.size();
                L.v.push_back({p0, p1, node});
                L.attr.push_back(0);
                L.alive.push_back(1);
                L.of_node[p0].push_back(id);
                L.of_node[p1].push_back(id);
                L.of_node[node].push_back(id);
                done = true;
            }
        }
    }
}

// The mortars' normals and surface Jacobians (SubdividedEdge_ConstructMortars /
// CurvedSubdividedEdge_ConstructMortars, QuadElement_EdgeMappings.incf:450-565) from the edge's
// n and dS as the linking left them: straight, n and dS / 2 at every mortar node; curved, dS n
// interpolated at 0.5 + 0.5 xi_N and 0.5 xi_S, its length halved and the normal its direction.
void mortar_geometry(const Edge &e, SubEdge &d) {
    const Spectral &s = spectral(e.N);
    for (int half = 0; half < 2; ++half) {
        const int NM = half == 0 ? d.NN : d.NS;
        const Spectral &sm = spectral(NM);
        std::vector<double> &nrm = half == 0 ? d.nN : d.nS, &dS = half == 0 ? d.dSN : d.dSS;
        nrm.assign(2 * (NM + 1), 0.0);
        dS.assign(NM + 1, 0.0);
        for (int p = 0; p <= NM; ++p) {
            if (!e.curved) {
                nrm[2 * p] = e.n[0]; nrm[2 * p + 1] = e.n[1];
                dS[p] = 0.5 * e.dS[0];
                continue;
            }
            const double x = half == 0 ? 0.5 + 0.5 * sm.xi[p] : 0.5 * sm.xi[p];
            double lj[kMaxP], v[2];
            lagrange_bary(x, e.N, s.xi, s.wb, lj);
            for (int c = 0; c < 2; ++c) {
                double acc = 0.0;
                for (int j = 0; j <= e.N; ++j) acc += (e.dS[j] * e.n[2 * j + c]) * lj[j];
                v[c] = acc;
            }
            dS[p] = 0.5 * std::sqrt(v[0] * v[0] + v[1] * v[1]);
            nrm[2 * p] = 0.5 * v[0] / dS[p];
            nrm[2 * p + 1] = 0.5 * v[1] / dS[p];
        }
    }
}

// SubdividedEdge_LinkWithElements (QuadElement_EdgeProcedures.incf:201-355) with
// SubdividedEdge_ComputeMortarsTransformationMatrices (:530-748).  The big element is LEFT and
// forward; the element with node 1 and the hanging node (quads[1] before linking) becomes NORTH
// when LEFT needs the edge inverted, SOUTH otherwise.
std::string link_subdivided(Mesh &m, Edge &e, int k) {
    int ids[3] = {e.nodes[0], e.nodes[1], e.hang};
    Elem &L = m.el[e.quads[0]];
    int pos, dir;
    locate(L.nodes, ids, pos, dir);
    if (pos < 0) return "subdivided edge not found in its element";
    L.edges[pos - 1] = k; L.dir[pos - 1] = 1; L.qpos[pos - 1] = 0; e.loc[0] = pos;
    int rn, rs;
    if (dir != 1) {
        rn = e.quads[1]; rs = e.quads[2];
        invert(e, pos == 1 || pos == 2);
        std::swap(ids[0], ids[1]);
    } else {
        rn = e.quads[2]; rs = e.quads[1];
        if (pos == 3 || pos == 4) {
            const int np = e.curved ? e.N + 1 : 1;
            for (int i = 0; i < 2 * np; ++i) e.n[i] = -e.n[i];
        }
    }
    e.quads[1] = rn; e.quads[2] = rs;
    e.sd = std::make_shared<SubEdge>();
    SubEdge &d = *e.sd;
    const int NL = L.N, NRN = m.el[rn].N, NRS = m.el[rs].N;
    d.NN = std::max(NL, NRN);
    d.NS = std::max(NL, NRS);
    mortar_geometry(e, d);
    const int sub_n[2] = {ids[2], ids[1]}, sub_s[2] = {ids[0], ids[2]};
    int dn, ds;
    Elem &RN = m.el[rn], &RS = m.el[rs];
    locate(RN.nodes, sub_n, pos, dn);
    if (pos < 0) return "subdivided edge not found in its north element";
    RN.edges[pos - 1] = k; RN.dir[pos - 1] = dn; RN.qpos[pos - 1] = 1; e.loc[1] = pos;
    locate(RS.nodes, sub_s, pos, ds);
    if (pos < 0) return "subdivided edge not found in its south element";
    RS.edges[pos - 1] = k; RS.dir[pos - 1] = ds; RS.qpos[pos - 1] = 2; e.loc[2] = pos;
    const Spectral &sL = spectral(NL), &sN = spectral(d.NN), &sS = spectral(d.NS), &sRN = spectral(NRN),
                   &sRS = spectral(NRS);
    const int PL = NL + 1, PN = d.NN + 1, PS = d.NS + 1, PRN = NRN + 1, PRS = NRS + 1;
    double xs[kMaxP], xn[kMaxP];
    for (int p = 0; p < PS; ++p) xs[p] = 0.5 * sS.xi[p];
    for (int p = 0; p < PN; ++p) xn[p] = 0.5 + 0.5 * sN.xi[p];
    d.TLS_F.assign(PS * PL, 0.0); interpolation_matrix(NL, d.NS, sL.xi, sL.wb, xs, d.TLS_F.data());
    d.TLN_F.assign(PN * PL, 0.0); interpolation_matrix(NL, d.NN, sL.xi, sL.wb, xn, d.TLN_F.data());
    d.TRS_F.assign(PS * PRS, 0.0); interpolation_matrix(NRS, d.NS, sRS.xi, sRS.wb, sS.xi, d.TRS_F.data());
    d.TRN_F.assign(PN * PRN, 0.0); interpolation_matrix(NRN, d.NN, sRN.xi, sRN.wb, sN.xi, d.TRN_F.data());
    // T_BKW = Minv T_FWD^T M with diagonal mass matrices: (1/w_el(i) T_FWD(l, i)) w_mortar(l)
    d.TLS_B.assign(PL * PS, 0.0); d.TLN_B.assign(PL * PN, 0.0);
    d.TRS_B.assign(PRS * PS, 0.0); d.TRN_B.assign(PRN * PN, 0.0);
    for (int i = 0; i < PL; ++i) {
        for (int l = 0; l < PS; ++l) d.TLS_B[i * PS + l] = ((1.0 / sL.w[i]) * d.TLS_F[l * PL + i]) * sS.w[l];
        for (int l = 0; l < PN; ++l) d.TLN_B[i * PN + l] = ((1.0 / sL.w[i]) * d.TLN_F[l * PL + i]) * sN.w[l];
    }
    for (int i = 0; i < PRS; ++i)
        for (int l = 0; l < PS; ++l) d.TRS_B[i * PS + l] = ((1.0 / sRS.w[i]) * d.TRS_F[l * PRS + i]) * sS.w[l];
    for (int i = 0; i < PRN; ++i)
        for (int l = 0; l < PN; ++l) d.TRN_B[i * PN + l] = ((1.0 / sRN.w[i]) * d.TRN_F[l * PRN + i]) * sN.w[l];
    // a small element running its half backwards: T_FWD's columns and T_BKW's rows reversed
    auto reverse = [](std::vector<double> &F, std::vector<double> &B, int Pm, int Pe) {
        for (int i = 0; i < Pm; ++i)
            for (int l = 0; l < Pe / 2; ++l) std::swap(F[i * Pe + l], F[i * Pe + Pe - 1 - l]);
        for (int i = 0; i < Pe / 2; ++i)
            for (int l = 0; l < Pm; ++l) std::swap(B[i * Pm + l], B[(Pe - 1 - i) * Pm + l]);
    };
    if (ds == -1) reverse(d.TRS_F, d.TRS_B, PS, PRS);
    if (dn == -1) reverse(d.TRN_F, d.TRN_B, PN, PRN);
    return "";
}

}  // namespace

const Spectral &spectral(int N) {
    static std::unique_ptr<Spectral> table[kMaxP];
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (!table[N]) {
        table[N].reset(new Spectral());
        make_spectral(*table[N], N);
    }
    return *table[N];
}

void mortar_matrices(int NLow, int N, double *Tf, double *Tb) {
    const Spectral &sl = spectral(NLow), &sh = spectral(N);
    const int Pl = NLow + 1, P = N + 1;
    interpolation_matrix(NLow, N, sl.xi, sl.wb, sh.xi, Tf);
    for (int i = 0; i < Pl; ++i)
        for (int l = 0; l < P; ++l) Tb[i * P + l] = ((1.0 / sl.w[i]) * Tf[l * Pl + i]) * sh.w[l];
}

double lagrange_at(const Spectral &s, double x, double *lj) {
    lagrange_bary(x, s.N, s.xi, s.wb, lj);
    return 0.0;
}

std::string build_mesh(const hg_mesh_desc &in, Mesh &m) {
    const int K = in.nelements;
    if (K <= 0 || in.nnodes <= 0 || !in.points || !in.quads || !in.order) return "empty mesh";
    m.pts.assign(in.points, in.points + 2 * (size_t)in.nnodes);
    for (int e = 0; e < K; ++e)
        for (int c = 0; c < 4; ++c)
            if (in.quads[4 * e + c] < 0 || in.quads[4 * e + c] >= in.nnodes) return "node id out of range";
    // faces in first-appearance order with their elements in element order
    std::unordered_map<uint64_t, int, PairHash> face_id;
    face_id.reserve((size_t)4 * K);
    std::vector<std::array<int, 2>> fnodes, fown;
    fnodes.reserve((size_t)2 * K + 16);
    fown.reserve((size_t)2 * K + 16);
    for (int e = 0; e < K; ++e)
        for (int f = 0; f < 4; ++f) {
            int a = in.quads[4 * e + f], b = in.quads[4 * e + (f + 1) % 4];
            auto it = face_id.find(key(a, b));
            int id;
            if (it == face_id.end()) {
                id = (int)fnodes.size();
                face_id.emplace(key(a, b), id);
                fnodes.push_back({a, b});
                fown.push_back({-1, -1});
            } else {
                id = it->second;
            }
            auto &o = fown[id];
            if (o[0] < 0) o[0] = e;
            else if (o[1] < 0 && o[0] != e) o[1] = e;
            else if (o[0] != e && o[1] != e) return "edge shared by more than two elements";
        }
    const int F = (int)fnodes.size();
    std::vector<int> fmark(F, 0);
    for (int b = 0; b < in.nbedges; ++b) {
        auto it = face_id.find(key(in.bedges[2 * b], in.bedges[2 * b + 1]));
        if (it == face_id.end()) return "boundary edge not found";
        fmark[it->second] = in.bmarker[b];
    }
    // subdivided edges: merged only where a node lies on exactly three faces, none of them on the
    // boundary (the merge's precondition), which conforming meshes rarely have
    std::vector<int> deg(in.nnodes, 0);
    std::vector<char> onb(in.nnodes, 0);
    for (int f = 0; f < F; ++f)
        for (int c = 0; c < 2; ++c) {
            ++deg[fnodes[f][c]];
            if (fmark[f]) onb[fnodes[f][c]] = 1;
        }
    bool candidates = false;
    for (int n = 0; n < in.nnodes && !candidates; ++n) candidates = deg[n] == 3 && !onb[n];
    EdgeList L;
    L.v.resize(F);
    L.attr.resize(F);
    L.alive.assign(F, 1);
    for (int f = 0; f < F; ++f) { L.v[f] = {fnodes[f][0], fnodes[f][1], -1}; L.attr[f] = fmark[f] != 0; }
    if (candidates) {
        L.of_node.resize(in.nnodes);
        for (int f = 0; f < F; ++f) {
            L.of_node[fnodes[f][0]].push_back(f);
            L.of_node[fnodes[f][1]].push_back(f);
            L.pair[key(fnodes[f][0], fnodes[f][1])].push_back(f);
        }
        merge_divided(in.nnodes, L);
    }
    // the edges in list order (computeElementOfEdges and searchDividedEdgeInElements,
    // MeshFile.f90:500-570,785-824: a subdivided edge's elements hold nodes 1 and 2, 1 and the
    // hanging node, 2 and the hanging node)
    m.ed.clear();
    m.ed.reserve(L.v.size());
    std::unordered_map<uint64_t, int, PairHash> edge_id;
    edge_id.reserve(L.v.size());
    for (size_t id = 0; id < L.v.size(); ++id) {
        if (!L.alive[id]) continue;
        Edge ed;
        ed.nodes[0] = L.v[id][0]; ed.nodes[1] = L.v[id][1];
        if (L.v[id][2] >= 0) {
            ed.hang = L.v[id][2];
            const uint64_t ks[3] = {key(ed.nodes[0], ed.nodes[1]), key(ed.nodes[0], ed.hang), key(ed.nodes[1], ed.hang)};
            for (int c = 0; c < 3; ++c) {
                auto it = face_id.find(ks[c]);
                if (it == face_id.end()) return "subdivided edge without its three elements";
                ed.quads[c] = fown[it->second][0];
            }
        } else {
            ed.quads[0] = fown[id][0];
            ed.quads[1] = fown[id][1];
            ed.bdry = ed.quads[1] < 0;
            ed.marker = fmark[id];
        }
        edge_id.emplace(key(ed.nodes[0], ed.nodes[1]), (int)m.ed.size());
        m.ed.push_back(std::move(ed));
    }
    face_id.swap(edge_id);
    const int E = (int)m.ed.size();
    std::vector<int> curve_of(E, -1);
    std::vector<double> cxy;
    if (in.ncurved > 0) {
        cxy.assign(in.cxy, in.cxy + (size_t)2 * in.np1 * in.ncurved);
        for (int c = 0; c < in.ncurved; ++c) {
            const int *cp = in.cedges + 2 * c;
            auto it = face_id.find(key(cp[0], cp[1]));
            if (it == face_id.end()) return "curved edge not found";
            curve_of[it->second] = c;
            const Edge &ed = m.ed[it->second];
            if (ed.nodes[0] == cp[1] && ed.nodes[1] == cp[0])
                for (int a = 0; a < in.np1 / 2; ++a)
                    for (int d = 0; d < 2; ++d)
                        std::swap(cxy[(c * in.np1 + a) * 2 + d], cxy[(c * in.np1 + in.np1 - 1 - a) * 2 + d]);
        }
    }
    m.el.assign(K, Elem());
    int64_t off = 0;
    for (int e = 0; e < K; ++e) {
        Elem &el = m.el[e];
        el.N = in.order[e]; el.P = el.N + 1;
        if (el.N < 1 || el.N > kMaxP - 3) return "polynomial order out of range";
        for (int c = 0; c < 4; ++c) { el.nodes[c] = in.quads[4 * e + c]; el.edges[c] = -1; }
        el.off = off; off += (int64_t)el.P * el.P;
    }
    m.nodes_total = off;
    for (int k = 0; k < E; ++k) {
        Edge &ed = m.ed[k];
        if (ed.hang >= 0) {
            // the big element's order; the curve at half the lowest of the three
            // (QuadMesh_Construct.incf:121-139)
            const int a = m.el[ed.quads[0]].N, b = m.el[ed.quads[1]].N, c = m.el[ed.quads[2]].N;
            const int lo = std::min(a, std::min(b, c));
            ed.N = a;
            ed.NCurve = lo > 1 ? lo / 2 : lo;
            // the reference sizes the NORTH mortar's reversal by the SOUTH element's order
            // (QuadElement_EdgeProcedures.incf:556-559): defined when the two agree
            if (b != c) return "subdivided edges need the two small elements of one order";
        } else if (ed.bdry) { ed.N = m.el[ed.quads[0]].N; ed.NCurve = ed.N; }
        else {
            // the higher order of the two elements; the curve at half the lower one
            // (QuadMesh_Construct.incf:104-118)
            const int a = m.el[ed.quads[0]].N, b = m.el[ed.quads[1]].N, lo = std::min(a, b);
            ed.N = std::max(a, b);
            ed.NCurve = lo > 1 ? lo / 2 : lo;
        }
        ed.curved = curve_of[k] >= 0;
        set_curve(m, ed, ed.curved ? &cxy[(size_t)2 * in.np1 * curve_of[k]] : nullptr, in.np1 - 1);
        if (ed.hang >= 0) {
            // the hanging node moves onto the edge at 0.5 (QuadMesh_Construct.incf:168-173)
            double X[2];
            point_at(m, ed, 0.5, X);
            m.pts[2 * ed.hang] = X[0];
            m.pts[2 * ed.hang + 1] = X[1];
        }
    }
    // straight conforming interior edges again, with the moved nodes (:177-185)
    for (int k = 0; k < E; ++k) {
        Edge &ed = m.ed[k];
        if (ed.hang < 0 && !ed.bdry && !ed.curved) set_curve(m, ed, nullptr, in.np1 - 1);
    }
    for (int k = 0; k < E; ++k) {
        Edge &ed = m.ed[k];
        if (ed.hang >= 0) {
            const std::string err = link_subdivided(m, ed, k);
            if (!err.empty()) return err;
            continue;
        }
        Elem &L = m.el[ed.quads[0]];
        int pos, dir;
        locate(L.nodes, ed.nodes, pos, dir);
        if (pos < 0) return "edge not found in its element";
        L.edges[pos - 1] = k; L.dir[pos - 1] = 1; L.qpos[pos - 1] = 0; ed.loc[0] = pos;
        if (dir != 1) invert(ed, pos == 1 || pos == 2);
        else if (pos == 3 || pos == 4) {
            int np = ed.curved ? ed.N + 1 : 1;
            for (int i = 0; i < 2 * np; ++i) ed.n[i] = -ed.n[i];
        }
        if (!ed.bdry) {
            Elem &R = m.el[ed.quads[1]];
            locate(R.nodes, ed.nodes, pos, dir);
            if (pos < 0) return "edge not found in its right element";
            R.edges[pos - 1] = k; R.dir[pos - 1] = dir; R.qpos[pos - 1] = 1; ed.loc[1] = pos;
            ed.inverse = dir != 1;
        }
    }
    for (auto &el : m.el) {
        for (int c = 0; c < 4; ++c)
            if (el.edges[c] < 0) return "element with an unlinked face";
        element_mappings(m, el);
    }
    for (auto &ed : m.ed) {
        const Spectral &s = spectral(ed.N);
        double A = 0.0;
        for (int i = 0; i <= ed.N; ++i) A += (ed.curved ? ed.dS[i] : ed.dS[0]) * s.w[i];
        ed.area = A;
        double v = m.el[ed.quads[0]].volume;
        if (!ed.bdry) v = std::fmin(v, m.el[ed.quads[1]].volume);
        if (ed.hang >= 0) v = std::fmin(v, m.el[ed.quads[2]].volume);
        ed.invh = A / v;
    }
    // zones (Mesh_ConstructZones, QuadMesh.f90:215-309)
    m.zones.assign(in.nmarkers, ZoneDesc());
    for (int z = 0; z < in.nmarkers; ++z) {
        ZoneDesc &Z = m.zones[z];
        Z.type = in.bc_type[z];
        Z.direction = in.bc_int[2 * z]; Z.partner = in.bc_int[2 * z + 1]; Z.isothermal = in.bc_int[2 * z];
        Z.nd_mode = in.bc_int[2 * z]; Z.nd_outflow = in.bc_int[2 * z + 1];
        for (int c = 0; c < 4; ++c) Z.q[c] = in.bc_real[4 * z + c];
        Z.Tw = in.bc_real[4 * z]; Z.vw[0] = in.bc_real[4 * z + 1]; Z.vw[1] = in.bc_real[4 * z + 2];
        Z.riemann = in.bc_riemann ? in.bc_riemann[z] : -1;
        switch (Z.type) {
        case HG_BC_PERIODIC: case HG_BC_DIRICHLET: case HG_BC_EULERWALL: case HG_BC_VISCOUSWALL:
        case HG_BC_FARFIELD: case HG_BC_PRESSUREOUTLET: case HG_BC_PRESSUREINLET: case HG_BC_NEWDIRICHLET: break;
        default: return "unsupported boundary condition type";
        }
    }
    for (int k = 0; k < E; ++k) {
        Edge &ed = m.ed[k];
        if (!ed.bdry) continue;
        // an element edge with no neighbour that is no boundary edge of the file, which the
        // merging of subdivided edges did not absorb
        if (ed.marker == 0) return "an edge with neither a neighbour nor a boundary marker";
        if (ed.marker < 1 || ed.marker > in.nmarkers) return "boundary edge without a valid marker";
        m.zones[ed.marker - 1].edges.push_back(k);
    }
    std::vector<char> zdone(in.nmarkers, 0);
    for (int z = 0; z < in.nmarkers; ++z) {
        ZoneDesc &Z1 = m.zones[z];
        if (Z1.type != HG_BC_PERIODIC || zdone[z]) continue;
        if (Z1.partner < 1 || Z1.partner > in.nmarkers) return "bad periodic partner marker";
        ZoneDesc &Z2 = m.zones[Z1.partner - 1];
        int coord = (Z1.direction == 1 && Z2.direction == 1) ? 1 : 0;
        for (int a : Z1.edges) {
            Edge &e1 = m.ed[a];
            if (e1.curved) return "curved periodic edges are not supported";
            for (int b : Z2.edges) {
                Edge &e2 = m.ed[b];
                if (e2.associated) continue;
                const double *p11 = &m.pts[2 * e1.nodes[0]], *p12 = &m.pts[2 * e1.nodes[1]];
                const double *p21 = &m.pts[2 * e2.nodes[0]], *p22 = &m.pts[2 * e2.nodes[1]];
                bool same = (almost_equal(p11[coord], p21[coord]) && almost_equal(p12[coord], p22[coord])) ||
                            (almost_equal(p11[coord], p22[coord]) && almost_equal(p12[coord], p21[coord]));
                if (!same) continue;
                double dot = e1.dX[0] * e2.dX[0] + e1.dX[1] * e2.dX[1];
                int inv = !(dot >= 0.0);
                e1.inverse = e2.inverse = inv;
                e1.associated = e2.associated = 1;
                e1.partner = b; e2.partner = a;
            }
        }
        zdone[z] = 1; zdone[Z1.partner - 1] = 1;
    }
    for (int k = 0; k < E; ++k) {
        const Edge &ed = m.ed[k];
        if (ed.bdry && m.zones[ed.marker - 1].type == HG_BC_PERIODIC && ed.partner < 0)
            return "periodic edge without a partner";
    }
    return "";
}

}  // namespace hg
