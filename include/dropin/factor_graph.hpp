/* Drop-in for ZeroSLAM's source/factor_graph.hpp.
 *
 * Put this directory ahead of ZeroSLAM's source/ for quoted includes (or copy
 * this file over source/factor_graph.hpp in a ZeroSLAM checkout): the
 * reference's callers of the factor graph, map::optimise (map.hpp:65-267) and
 * through it slam::process_frame (slam.hpp:316, 387, 469, 472), then build and
 * solve their bundle adjustment on the GPU engine with their text unchanged.
 *
 * It provides what those callers take from the reference header:
 *   - the factor_graph namespace (factor_graph, vertex_pose, vertex_point_xyz,
 *     edge_reprojection<camera>, the four losses; include/zsba_factor_graph.hpp)
 *     with ZeroSLAM's matrix::matrix<double, 0, 0> as the parameter type
 *     (map.hpp:136-137, 158, 240-250);
 *   - the headers the reference header brings in and map.hpp relies on:
 *     lie.hpp (lie::se3, lie::so3, map.hpp:136, 242-243) with math.hpp
 *     (math::sqrt, map.hpp:75) and matrix.hpp.
 * The include guard is the reference header's, so exactly one of the two is
 * ever in a translation unit.  The engine precision of every graph defaults to
 * ZSBA_SHIM_PRECISION (ZSBA_PRECISION_FP64, the reference's arithmetic, when
 * not defined).  The library is zeroslam_amd/libzsba.so (include/zsba.h).
 */
#pragma once
#ifndef FACTOR_GRAPH_HPP
#define FACTOR_GRAPH_HPP

#include "lie.hpp"
#include "matrix.hpp"

#ifndef ZSBA_SHIM_MATRIX
#define ZSBA_SHIM_MATRIX ::matrix::matrix<double, 0, 0>
#endif
#include "zsba_factor_graph.hpp"

namespace factor_graph = ::zsba::factor_graph;

#endif  // FACTOR_GRAPH_HPP
