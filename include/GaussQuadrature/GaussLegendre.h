// Cubature rules on the unit tetrahedron and unit triangle, and the
// order -> rule selector used by the DG scheme.
// Names and selector mapping follow the reference's
// include/GaussQuadrature/GaussLegendre.h:22-111 and AutoQuadSelectorTri.h:1-22;
// the point/weight doubles come from quad_rules.inc (generated, see
// tests/golden/gen_quad_rules.py).  Tet weights sum to 1/6, triangle weights to 1/2.
#pragma once
#include "base/Type.h"
#include "GaussQuadrature/quad_rules.inc"

template <uInt Order, typename DefaultQuad>
struct AutoQuadSelector {
    using type = DefaultQuad;
};

#define FVDG_DEFINE_RULE(Name, TAG, DIM)                                                   \
    struct Name {                                                                          \
        static constexpr int num_points = TAG##_N;                                         \
        static constexpr std::array<std::array<Scalar, DIM>, num_points> points = {TAG##_PTS}; \
        static constexpr std::array<Scalar, num_points> weights = TAG##_W;                 \
    };

namespace GaussLegendreTet {
struct Auto {
    static constexpr int num_points = 0;
};
FVDG_DEFINE_RULE(Degree5Points15, FVDG_TET_DEGREE5POINTS15, 3)
FVDG_DEFINE_RULE(Degree7Points31, FVDG_TET_DEGREE7POINTS31, 3)
FVDG_DEFINE_RULE(Degree9Points57, FVDG_TET_DEGREE9POINTS57, 3)
FVDG_DEFINE_RULE(Degree11Points94, FVDG_TET_DEGREE11POINTS94, 3)
}  // namespace GaussLegendreTet

namespace GaussLegendreTri {
struct Auto {
    static constexpr int num_points = 0;
};
FVDG_DEFINE_RULE(Degree4Points7, FVDG_TRI_DEGREE4POINTS7, 2)
FVDG_DEFINE_RULE(Degree10Points24, FVDG_TRI_DEGREE10POINTS24, 2)
FVDG_DEFINE_RULE(Degree13Points36, FVDG_TRI_DEGREE13POINTS36, 2)
}  // namespace GaussLegendreTri
#undef FVDG_DEFINE_RULE

// Order -> rule (reference GaussLegendre.h:84-111 for tets, AutoQuadSelectorTri.h:1-22 for faces).
template <> struct AutoQuadSelector<0, GaussLegendreTet::Auto> { using type = GaussLegendreTet::Degree5Points15; };
template <> struct AutoQuadSelector<1, GaussLegendreTet::Auto> { using type = GaussLegendreTet::Degree5Points15; };
template <> struct AutoQuadSelector<2, GaussLegendreTet::Auto> { using type = GaussLegendreTet::Degree5Points15; };
template <> struct AutoQuadSelector<3, GaussLegendreTet::Auto> { using type = GaussLegendreTet::Degree7Points31; };
template <> struct AutoQuadSelector<4, GaussLegendreTet::Auto> { using type = GaussLegendreTet::Degree9Points57; };
template <> struct AutoQuadSelector<5, GaussLegendreTet::Auto> { using type = GaussLegendreTet::Degree11Points94; };
template <> struct AutoQuadSelector<0, GaussLegendreTri::Auto> { using type = GaussLegendreTri::Degree4Points7; };
template <> struct AutoQuadSelector<1, GaussLegendreTri::Auto> { using type = GaussLegendreTri::Degree4Points7; };
template <> struct AutoQuadSelector<2, GaussLegendreTri::Auto> { using type = GaussLegendreTri::Degree10Points24; };
template <> struct AutoQuadSelector<3, GaussLegendreTri::Auto> { using type = GaussLegendreTri::Degree10Points24; };
template <> struct AutoQuadSelector<4, GaussLegendreTri::Auto> { using type = GaussLegendreTri::Degree10Points24; };
template <> struct AutoQuadSelector<5, GaussLegendreTri::Auto> { using type = GaussLegendreTri::Degree13Points36; };

// Run-time view of the same rules, for code that selects the order at run time.
struct QuadRuleView {
    int num_points;
    int dim;
    const Scalar* points;  // num_points * dim, row-major
    const Scalar* weights;
};

template <typename Rule>
inline QuadRuleView quad_rule_view() {
    return {Rule::num_points, (int)Rule::points[0].size(), Rule::points[0].data(), Rule::weights.data()};
}

inline QuadRuleView tet_rule_for_order(int order) {
    switch (order) {
        case 0: case 1: case 2: return quad_rule_view<GaussLegendreTet::Degree5Points15>();
        case 3: return quad_rule_view<GaussLegendreTet::Degree7Points31>();
        case 4: return quad_rule_view<GaussLegendreTet::Degree9Points57>();
        case 5: return quad_rule_view<GaussLegendreTet::Degree11Points94>();
        default: throw std::runtime_error("no tetrahedron rule for order " + std::to_string(order));
    }
}
inline QuadRuleView tri_rule_for_order(int order) {
    switch (order) {
        case 0: case 1: return quad_rule_view<GaussLegendreTri::Degree4Points7>();
        case 2: case 3: case 4: return quad_rule_view<GaussLegendreTri::Degree10Points24>();
        case 5: return quad_rule_view<GaussLegendreTri::Degree13Points36>();
        default: throw std::runtime_error("no triangle rule for order " + std::to_string(order));
    }
}
