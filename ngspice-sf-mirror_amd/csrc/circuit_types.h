/*
 * The engine's circuit description (the counterpart of the netlist ->
 * CKTcircuit build): nodes 1..nnodes (0 = ground), BSIM4 instances,
 * capacitors, resistors, independent voltage sources, .ic values and the
 * options; shared by the batched transient and the one-large-circuit
 * transient.
 */
#ifndef NGMI_CIRCUIT_TYPES_H
#define NGMI_CIRCUIT_TYPES_H

#include <array>
#include <string>
#include <vector>
#include "ngmi_internal.h"

struct ngmi_circuit {
    int nnodes = 0;
    struct Mos {
        ngmi_bsim4_model *model;
        int node[4];
        std::vector<std::string> names;
        std::vector<double> values;
    };
    std::vector<Mos> mos;
    std::vector<std::array<double, 3>> caps;   /* p n C */
    std::vector<std::array<double, 3>> res;    /* p n R */
    struct Vs {
        int p, n;
        double dc;
        int npulse;
        double pulse[7];
    };
    std::vector<Vs> vs;
    std::vector<double> ic;
    std::vector<char> ic_given;
    double reltol = 1e-3, abstol = 1e-12, vntol = 1e-6, chgtol = 1e-14, trtol = 7.0, gmin = 1e-12;
    int method = 1, maxord = 2;
    double pivrel = 1e-3;   /* CKTpivotRelTol, KLU's pivot tolerance */
    /* optran's step and final time (ngspice's default 'optran 1 1 1 100n
     * 10u 0', init.c:77-93); opstep 0 deselects it */
    double opstep = 1e-7, opfinal = 1e-5;
    int uic = 0;   /* .tran ... uic: MODEUIC, the .ic values as the initial state */
    /* CKTop options (cktntask.c:94-132 defaults) */
    int noopiter = 0, gminsteps = 1, srcsteps = 1, itl2 = 50;
    double gminfactor = 10.0, gshunt = 0.0;
    /* optional pivot sequence (1-based unknowns); empty = the engine's own,
     * or KLU's (restated klu_analyze on the circuit's pattern) when
     * klu_order is set */
    std::vector<int> order_rows, order_cols;
    int klu_order = 0;
};

#endif
