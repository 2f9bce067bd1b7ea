// Host-side mesh model of the HORSES2D DGSEM path: spectral operators, connectivity
// with the reference's numbering, element/edge geometry and boundary zones.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "horses2d_gpu.h"

namespace hg {

constexpr int kMaxP = 16;  // polynomial orders 1..14 on the host side

// Legendre-Gauss spectral operators on the reference element [0,1]
// (Approximation/NodesAndWeightsClass.f90:251-332).
struct Spectral {
    int N = 0, P = 0;
    double xi[kMaxP], w[kMaxP], wb[kMaxP];
    double D[kMaxP][kMaxP];     // D[i][j] = l_j'(x_i)
    double hatD[kMaxP][kMaxP];  // hatD(i,l) = w_l D(l,i) / w_i
    double lb[2][kMaxP];        // [0]: l_j(0) (reference LEFT), [1]: l_j(1) (reference RIGHT)
    double lbw[2][kMaxP], dlb[2][kMaxP], dlbw[2][kMaxP];
};
const Spectral &spectral(int N);
double lagrange_at(const Spectral &s, double x, double *lj);  // barycentric l_j(x)
// p-non-conforming edge of orders NLow < N (Edge_LinkWithElements, QuadElement_EdgeProcedures.incf:
// 105-124): T_forward [(N+1) x (NLow+1)] (PolynomialInterpolationMatrix from the low order's
// nodes to the edge's) and T_backward [(NLow+1) x (N+1)] = Minv_low T_forward^T M_edge
// (TripleMatrixProduct with diagonal mass matrices: (1/w_low(i) T_forward(l, i)) w_edge(l)).
void mortar_matrices(int NLow, int N, double *Tf, double *Tb);

// A subdivided edge's two mortars (SubdividedEdge_ConstructMortars and
// SubdividedEdge_ComputeMortarsTransformationMatrices, QuadElement_EdgeMappings.incf:450-565,
// QuadElement_EdgeProcedures.incf:530-748): NORTH runs from the hanging node to node 2, SOUTH
// from node 1 to the hanging node; orders NN = max(N_left, N_north), NS = max(N_left, N_south).
// Matrices row-major: T_L?_F [mortar][left], T_L?_B [left][mortar], T_R?_F [mortar][small],
// T_R?_B [small][mortar] (a small element running its half backwards has them reversed).
// Normals and surface Jacobians per mortar node (straight: the edge's n and dS / 2 at all).
struct SubEdge {
    int NN = 0, NS = 0;
    std::vector<double> TLN_F, TLS_F, TLN_B, TLS_B, TRN_F, TRS_F, TRN_B, TRS_B;
    std::vector<double> nN, nS, dSN, dSS;
};

struct Edge {
    int nodes[2] = {-1, -1};   // after the linking inversions
    int quads[3] = {-1, -1, -1};   // LEFT (lowest id) / boundary element, RIGHT (subdivided: NORTH, SOUTH)
    int loc[3] = {0, 0, 0};        // 1..4 = bottom, right, top, left
    int inverse = 0, marker = 0, bdry = 0, curved = 0;
    int N = 0, NCurve = 0;
    int partner = -1, associated = 0;
    int hang = -1;                 // subdivided edge (MeshFile.f90:364-490): the hanging node
    std::shared_ptr<SubEdge> sd;
    std::vector<double> X, dX, n, dS;  // straight: dX, n, dS use entry 0
    double area = 0.0, invh = 0.0;
};

struct Elem {
    int N = 0, P = 0;
    int nodes[4];
    int edges[4];   // per position 1..4
    int dir[4];     // edgesDirection (+1 forward / -1 backward)
    int qpos[4];    // 0: LEFT/boundary side of the edge, 1: RIGHT (subdivided: NORTH) side, 2: SOUTH
    std::vector<double> X;    // [(i + P j)*2 + d]
    std::vector<double> Ja;   // [(r*2+c)*P*P + i + P j]
    std::vector<double> jac;  // [i + P j]
    double volume = 0.0;
    int64_t off = 0;
};

struct ZoneDesc {
    int type = 0, direction = 0, partner = 0, isothermal = 0, riemann = -1;   // isothermal: also the outflow type
    int nd_mode = 0, nd_outflow = 0;   // newDirichlet: HG_NEWDIRICHLET_* mode and outflow pressure
    double q[4] = {0, 0, 0, 0};
    double Tw = 1.0, vw[2] = {0, 0};
    std::vector<int> edges;
};

struct Mesh {
    std::vector<double> pts;
    std::vector<Elem> el;
    std::vector<Edge> ed;
    std::vector<ZoneDesc> zones;
    int64_t nodes_total = 0;
};

// Builds the mesh; returns an empty string on success or an error message.
std::string build_mesh(const hg_mesh_desc &in, Mesh &m);

}  // namespace hg
