// performdocking.h -- the MI355X backend's docking driver.  Same declarations
// as the reference's host/inc/performdocking.h.Cuda:65-96, so that the
// reference's main.cpp builds against it unchanged.
#ifndef PERFORMDOCKING_H_
#define PERFORMDOCKING_H_

#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <string>
#include <vector>

#include "profile.hpp"
#include "processgrid.h"
#include "miscellaneous.h"
#include "processligand.h"
#include "getparameters.h"
#include "calcenergy.h"
#include "processresult.h"
#include "simulation_state.hpp"
#include "GpuData.h"

std::vector<int> get_gpu_pool();

void setup_gpu_for_docking(
                           GpuData&     cData,
                           GpuTempData& tData
                          );

void finish_gpu_from_docking(
                             GpuData&     cData,
                             GpuTempData& tData
                            );

int docking_with_gpu(
                     const Gridinfo*        mygrid,
                           Dockpars*        mypars,
                     const Liganddata*      myligand_init,
                     const Liganddata*      myxrayligand,
                           Profile&         profile,
                     const int*             argc,
                           char**           argv,
                           SimulationState& sim_state,
                           GpuData&         cData,
                           GpuTempData&     tData,
                           std::string*     output = NULL
                    );

double check_progress(
                      int*           evals_of_runs,
                      int            generation_cnt,
                      int            max_num_of_evals,
                      int            max_num_of_gens,
                      int            num_of_runs,
                      unsigned long& total_evals
                     );

#endif /* PERFORMDOCKING_H_ */
