// Umbrella header (reference: include/Mesh/Mesh.h).
#pragma once
#include "Mesh/GeneralMesh.h"
#include "Mesh/OrthHexMesh.h"
#include "Mesh/ComputingMesh.h"
