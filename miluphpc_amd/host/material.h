// material.h -- one material, field names of the reference's Material class
// (include/materials/material.cuh:30-160): interactions = target neighbour count,
// artificialViscosity {alpha, beta}, eos {type, polytropic_K, polytropic_gamma}.
// The reference reads these from a libconfig file (material_handler.cpp:28); here the
// caller fills the struct.
#pragma once

#include "parameter.h"

namespace EquationOfStates {
enum Type {
    EOS_TYPE_POLYTROPIC_GAS = 0,
    EOS_TYPE_ISOTHERMAL_GAS = 3,
    EOS_TYPE_IDEAL_GAS = 9,
    EOS_TYPE_LOCALLY_ISOTHERMAL_GAS = 12,
    EOS_TYPE_ISOTHERMAL_CONST_CS = 100   // build extension: p = cs^2 rho, cs given
};
}

struct ArtificialViscosity {
    real alpha = 1.0;
    real beta = 2.0;
};

struct EqOfSt {
    int type = EquationOfStates::EOS_TYPE_IDEAL_GAS;
    real polytropic_K = 0.0;
    real polytropic_gamma = 5.0 / 3.0;
    real cs = 0.0;   // sound speed of EOS_TYPE_ISOTHERMAL_CONST_CS
};

struct Material {
    integer ID = 0;
    integer interactions = 64;
    real sml = 0.0;
    ArtificialViscosity artificialViscosity;
    EqOfSt eos;
    // accepted neighbour-count band of the plain h-iteration
    integer bandLow = 50;
    integer bandHigh = 80;
};
