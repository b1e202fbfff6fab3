// ORACLE — test infrastructure only (see oracle_interval.hpp).
// Scene state shared by the oracle's broad phase, constraint set and
// assembly: the data RigidBodyAssembler holds
// (src/physics/rigid_body_assembler.cpp:19-118, src/physics/rigid_body.cpp:124-142).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "oracle_ccd.hpp"

namespace oracle {

struct Scene {
    int nb = 0, nv = 0, ne = 0, nf = 0;
    std::vector<double> V;   // [nv][3] body-local
    std::vector<int> E;      // [ne][2]
    std::vector<int> F;      // [nf][3]
    std::vector<int> FE;     // [nf][3]
    std::vector<int> bv, be, bf;  // [nb+1] element ranges per body
    std::vector<int> group;  // [nb]
    std::vector<uint8_t> is_static;
    // derived
    std::vector<int> v_body, e_body, f_body;
    std::vector<double> r_max, edge_len;
    double average_edge_length = 0;
};

Scene& scene();
void scene_derive(Scene& s);

// RigidBody::world_vertices (rigid_body.tpp:11-16) for one body.
void world_vertices(const Scene& s, int b, const double pose[6], std::vector<double>& out);

struct Candidates {
    std::vector<std::array<int, 2>> ee;  // (e0 < e1)
    std::vector<std::array<int, 2>> fv;  // (face, vertex)
};

// detect_collision_candidates_rigid with HASH_GRID, continuous
// (src/ccd/rigid/broad_phase.cpp:127-203) and discrete (:19-90).
void broadphase_ccd(const Scene& s, const double* p0, const double* p1, double inflation_radius,
                    Candidates& out);
void broadphase_static(const Scene& s, const double* poses, double inflation_radius, Candidates& out);

// detect_collision_candidates_rigid with BVH (src/ccd/rigid/broad_phase.cpp:92-121,
// 205-230; oracle_bvh.cpp): continuous over poses p0 -> p1, or discrete at
// p0. EE candidates keep the reference's orientation (an edge of the body
// with fewer faces first); both lists sorted. bvh_supported: every vertex
// lies on an edge and every edge on a face (no codimensional elements).
void broadphase_bvh(const Scene& s, const double* p0, const double* p1, double inflation_radius, bool continuous,
                    Candidates& out);
std::vector<std::array<int, 2>> close_bodies_bvh(const Scene& s, const double* p0, const double* p1, double r);
bool bvh_supported(const Scene& s);

// [nv][7] included flag, AABB min, AABB max of the last broad-phase call.
std::vector<double>& last_vertex_aabbs();

// Exact orient3d sign and the edge-triangle rule; every intersecting
// (edge, face) pair of the close body pairs, sorted
// (RigidBodyProblem::detect_intersections, rigid_body_problem.cpp:289-360).
int orient3d_sign(const double* a, const double* b, const double* c, const double* d);
bool edge_triangle_intersect(const double* p, const double* q, const double* t0, const double* t1, const double* t2);
std::vector<std::array<int, 2>> detect_intersections(const Scene& s, const double* poses);

}  // namespace oracle
