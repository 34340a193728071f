// The device-resident B^T diag(d) A plan behind include/sdl_device.h.
#pragma once

#include "spgemm.h"

struct sdl_btda_plan
{
    int m = 0, nA = 0, nB = 0;
    sdl::DBuf<int> Ap, Ai;            // A's structure (copied)
    sdl::DBuf<int> Lp, Li, Lperm;     // B's columns: CSC structure + CSR positions
    sdl::ProductPlan P;
};

// Plan creation with an explicit symbolic mode and predictor (may be null).
int sdl_btda_plan_create_mode(int m, int nA, int nB, const int *dAp, const int *dAi,
                              const int *dBp, const int *dBi, sdl::SymMode mode,
                              const sdl::RowPredictor *pred, sdl_btda_plan **out,
                              bool force_pull = false, bool rank_refill = false);
