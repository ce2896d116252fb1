// Application-defined exact-solution hooks (reference: include/base/exact.h:6-52).
//
// The residual's Dirichlet faces take their boundary state from these functions,
// evaluated at reference point Xi of the face's side-0 cell and time t
// (reference use: src/DG/DG_Schemes/ExplicitConvection.cpp:47-52).  As in the
// reference, they are declared here and defined by the application; a program
// that instantiates ExplicitConvection<...>::eval or eval_boundarys must define
// the five conserved-variable hooks rho_Xi .. rhoe_Xi (or define
// FVDG_NO_EXACT_HOOKS before including the scheme headers and supply boundary
// states through ExplicitConvection::set_boundary_function).
#pragma once
#include "Mesh/ComputingMesh.h"
#include "base/Type.h"

#define FVDG_FIELD_HOOKS(name)                                                  \
    Scalar name##_xyz(const vector3f& xyz, Scalar t = 0.0);                     \
    Scalar name##_Xi(const CompTetrahedron& cell, const vector3f& Xi, Scalar t = 0.0); \
    Scalar rho##name##_Xi(const CompTetrahedron& cell, const vector3f& Xi, Scalar t = 0.0);

FVDG_FIELD_HOOKS(rho)
FVDG_FIELD_HOOKS(u)
FVDG_FIELD_HOOKS(v)
FVDG_FIELD_HOOKS(w)
FVDG_FIELD_HOOKS(p)
FVDG_FIELD_HOOKS(e)
#undef FVDG_FIELD_HOOKS
