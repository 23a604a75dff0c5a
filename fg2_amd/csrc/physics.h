// physics.h -- the physics modules this library is built with, and the
// dispatch the kernel launchers and the host driver derive from that list.
//
// A module is one header under modules/ holding one struct: the counterpart
// of the reference's src/<module>/ directory, which its Makefile selects at
// build time (Makefile:19-22).  Here every listed module is built in and a
// run picks one by name.  A module struct provides
//   NAME          the name runs are set up with ("hydro", ...)
//   FIELDS        the state variables, comma-separated                (state.h)
//   NV            number of variables, NVAR = sizeof(S)/sizeof(R)   (src/fg2.h:47-48, state.h)
//   ODD           variable negated by the reflecting dim-2 boundary (bcond.cuh)
//   ND, D12V      variables whose mixed derivative D12 the RHS uses
//   FLOP          eqns_flop of the reference's perf model          (config.h)
//   PARAMS        parameter names and defaults, in Params::par order (eqns.cuh)
//   settings()    the geometry the module imposes                   (config.h)
//   ONE_PASS      whether the one-pass step kernel applies (register rings fit)
//   ROW_TABLE     whether it reads a per-row table (row_entry(), Params::rowtab)
//   rhs()         the module's eqns()                               (eqns.cuh)
//   DERIV_FIRST   whether rhs() also comes in two parts, deriv_var() /
//                 deriv_mixed() (the derivatives of one variable, the
//                 mixed ones) and rhs_d() (the rest, from those), so that
//                 the substage kernel keeps derivatives, not stencil
//                 primitives, live through the right-hand side
//   STEP_ONE_BARRIER  the one-pass step's schedule on periodic grids and
//                 slabs: one workgroup barrier per row (stages lagged by four
//                 rows, double-buffered LDS rows) or one per stage (fewer
//                 registers); the bounded kernel always takes one per row
//   first_order() and first_order_on(), HAS_FIRST_ORDER, RANDOM_PHASES
//                 the point-wise pre-step, when it runs, and whether it
//                 draws two random phases per step                  (Euler.cuh)
//   lam_adv(), lam_diff(), CELL_DIFF, diffusivity()
//                 the CFL rates: per-cell advective and diffusive rates, or
//                 a constant D_max from the parameters
//   sums()        FP64 mass and momentum of one cell (fg2_sums)
//   ic_of(), ic() the named initial conditions and their cell function (init.h)
// Adding a module is adding its header and its struct to Modules below:
// nothing else in the library names a module.
#pragma once
#include <cstring>
#include "modules/hydro.h"
#include "modules/hydro_isot.h"
#include "modules/disk.h"

namespace fg2 {

template<class... M> struct ModuleList { static constexpr int size = sizeof...(M); };

// The modules, in the order of their ids.
using Modules = ModuleList<Hydro, HydroIsot, Disk>;

// f(M{}) for the module with this id (hipErrorInvalidValue for no module).
template<class F, class... M>
hipError_t dispatch_module(ModuleList<M...>, int id, F &&f)
{
  hipError_t out = hipErrorInvalidValue;
  int k = 0;
  (void)((k++ == id ? (out = f(M{}), true) : false) || ...);
  return out;
}
template<class F> hipError_t by_module(int id, F &&f) { return dispatch_module(Modules{}, id, f); }

// What the host driver needs to know of a module, from its struct.
struct ModuleInfo {
  const char *name, *fields;
  int nv, odd, flop, npar;
  bool one_pass, row_table, has_first_order, random_phases;
  const ParamDef *params;
  void (*settings)(Geometry &, int, int);
  bool (*first_order_on)(const float *);
  double (*diffusivity)(const double *);
  int (*ic_of)(const char *);
};

template<class M> ModuleInfo module_info_of()
{
  return ModuleInfo{M::NAME, M::FIELDS, M::NV, M::ODD, M::FLOP,
                    (int)(sizeof(M::PARAMS) / sizeof(M::PARAMS[0])),
                    M::ONE_PASS, M::ROW_TABLE, M::HAS_FIRST_ORDER, M::RANDOM_PHASES,
                    M::PARAMS, &M::settings, &M::first_order_on, &M::diffusivity, &M::ic_of};
}

template<class... M> const ModuleInfo *module_table(ModuleList<M...>)
{
  static const ModuleInfo table[] = {module_info_of<M>()...};
  return table;
}

inline int module_count() { return Modules::size; }
// id is one of [0, module_count())
inline const ModuleInfo &module_info(int id) { return module_table(Modules{})[id]; }
// The id of the module of this name, or -1.
inline int module_id(const char *name)
{
  if(!name) return -1;
  for(int k = 0; k < module_count(); ++k)
    if(!strcmp(name, module_info(k).name)) return k;
  return -1;
}
inline int nvar_of(int id) { return module_info(id).nv; }

} // namespace fg2
