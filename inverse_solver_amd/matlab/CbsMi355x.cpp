// MATLAB MEX gateway onto libcbs_hip.so (include/cbs_hip.h).
//
// Keeps the action codes of the reference's CudaBorn MEX (Codes/CUDA_CODE/main.cpp:53-213)
// so that FORWARD_SOLVER_CONVERGENT_BORN_MI355X.m (beside this file) mirrors
// FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m with the MEX name changed:
//   id = CbsMi355x(uint32(0), uint32(0), geometry_struct, refocusing_util, kernel_trans, kernel_ref)
//        CbsMi355x(uint32(1), id, potential_struct, V)
//   u  = CbsMi355x(uint32(2), id, u_in)                    (3D incident field, RI grid, n_rhs on dim 4)
//        CbsMi355x(uint32(3), id)
//   [t, r, u3] = CbsMi355x(uint32(4), id, source_2d)       (filtered real-space 2D fields, n_rhs on dim 3)
// Action 0 takes either CudaBorn's first octants (floor(conv/2)+1 x floor(conv/2)+1 x
// floor(RI_z/2)+1 of refocusing_util and kernel_trans, as FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m:28-38
// passes them; expanded by cbs_set_io_kernels_octant with the reference's symmetry) or the
// whole (conv_x, conv_y, RI_z) kernels plus kernel_ref.  Differences from CudaBorn: one
// handle runs on one device (geometry.used_gpu(1)), and all angles of a call are solved as
// one batch.
//
// Build (with MATLAB R2018a+):  mex -R2018a CbsMi355x.cpp -I../../include -L.. -lcbs_hip
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "cbs_hip.h"
#include "mex.h"

namespace {

double num_at(const mxArray* a, size_t i) {
  if (i >= mxGetNumberOfElements(a)) mexErrMsgIdAndTxt("cbs:arg", "numeric field too short");
  switch (mxGetClassID(a)) {
    case mxDOUBLE_CLASS: return mxGetDoubles(a)[i];
    case mxSINGLE_CLASS: return mxGetSingles(a)[i];
    case mxUINT32_CLASS: return mxGetUint32s(a)[i];
    case mxINT32_CLASS: return mxGetInt32s(a)[i];
    case mxUINT8_CLASS: return mxGetUint8s(a)[i];
    case mxLOGICAL_CLASS: return mxGetLogicals(a)[i];
    default: mexErrMsgIdAndTxt("cbs:arg", "unsupported numeric class in a parameter field");
  }
  return 0;
}

const mxArray* field(const mxArray* s, const char* name) {
  const mxArray* f = mxGetField(s, 0, name);
  if (!f) mexErrMsgIdAndTxt("cbs:arg", (std::string("missing parameter field: ") + name).c_str());
  return f;
}

const mxComplexSingle* complex_single(const mxArray* a, size_t expect, const char* what) {
  if (!mxIsSingle(a) || !mxIsComplex(a))
    mexErrMsgIdAndTxt("cbs:arg", (std::string(what) + " must be complex single").c_str());
  if (expect && mxGetNumberOfElements(a) != expect)
    mexErrMsgIdAndTxt("cbs:arg", (std::string(what) + " has the wrong number of elements").c_str());
  return mxGetComplexSingles(a);
}

void check(int status) {
  if (status != CBS_OK) mexErrMsgIdAndTxt("cbs:lib", "%s", cbs_last_error());
}

struct Entry {
  cbs_handle h;
  cbs_geometry geo;
};
std::vector<Entry*> g_handles;   // index = id

Entry* entry(const mxArray* id) {
  const uint32_t i = (uint32_t)num_at(id, 0);
  if (i >= g_handles.size() || !g_handles[i]) mexErrMsgIdAndTxt("cbs:id", "Id not found");
  return g_handles[i];
}

void at_exit() {
  for (Entry* e : g_handles)
    if (e) {
      cbs_destroy(e->h);
      delete e;
    }
  g_handles.clear();
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static bool registered = false;
  if (!registered) {
    mexAtExit(at_exit);
    registered = true;
  }
  if (nrhs < 2) mexErrMsgIdAndTxt("cbs:arg", "Need at least the action and id");
  if (!mxIsUint32(prhs[0]) || !mxIsUint32(prhs[1])) mexErrMsgIdAndTxt("cbs:arg", "action and id must be uint32");
  const uint32_t action = (uint32_t)num_at(prhs[0], 0);
  switch (action) {
    case 0: {  // create + I/O kernels
      if (nrhs < 5) mexErrMsgIdAndTxt("cbs:arg", "Need the parameters and the refocusing/transmission kernels");
      const mxArray* s = prhs[2];
      Entry* e = new Entry();
      for (int j = 0; j < 3; ++j) {
        e->geo.simulation_size[j] = (uint32_t)num_at(field(s, "simulation_size"), j);
        e->geo.RI_size[j] = (uint32_t)num_at(field(s, "RI_size"), j);
        e->geo.field_size[j] = (uint32_t)num_at(field(s, "field_size"), j);
        e->geo.conv_size[j] = (uint32_t)num_at(field(s, "conv_size"), j);
      }
      const mxArray* gpu = mxGetField(s, 0, "used_gpu");
      e->geo.device = (gpu && mxGetNumberOfElements(gpu)) ? (int32_t)num_at(gpu, 0) : 0;
      const mxArray* rc = mxGetField(s, 0, "RI_center");
      e->geo.RI_center[0] = rc ? (int32_t)num_at(rc, 0) : 0;
      e->geo.RI_center[1] = rc ? (int32_t)num_at(rc, 1) : 0;
      const mxArray* pn = mxGetField(s, 0, "pole_num");   // 3: vector_simulation (:294-297)
      e->geo.pole_num = (pn && mxGetNumberOfElements(pn)) ? (int32_t)num_at(pn, 0) : 1;
      check(cbs_create(&e->geo, &e->h));
      const size_t nk = (size_t)e->geo.conv_size[0] * e->geo.conv_size[1] * e->geo.RI_size[2];
      const size_t no = (size_t)(e->geo.conv_size[0] / 2 + 1) * (e->geo.conv_size[1] / 2 + 1) * (e->geo.RI_size[2] / 2 + 1);
      if (nrhs == 5 && mxGetNumberOfElements(prhs[3]) == no && no != nk) {
        check(cbs_set_io_kernels_octant(e->h, complex_single(prhs[3], no, "green_for_cuda"),
                                        complex_single(prhs[4], no, "green_for_cuda_2"), CBS_MEM_HOST));
      } else {
        const void* kref = nrhs > 5 ? (const void*)complex_single(prhs[5], nk, "kernel_ref") : nullptr;
        check(cbs_set_io_kernels(e->h, complex_single(prhs[3], nk, "refocusing_util"),
                                 complex_single(prhs[4], nk, "kernel_trans"), kref, CBS_MEM_HOST));
      }
      g_handles.push_back(e);
      if (nlhs > 0) {
        plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT32_CLASS, mxREAL);
        mxGetUint32s(plhs[0])[0] = (uint32_t)(g_handles.size() - 1);
      }
      break;
    }
    case 1: {  // potential
      if (nrhs < 4) mexErrMsgIdAndTxt("cbs:arg", "Need to give both parameters and scattering potential");
      Entry* e = entry(prhs[1]);
      const mxArray* s = prhs[2];
      cbs_potential_params p{};
      p.k0_nm = num_at(field(s, "k0_nm"), 0);
      p.eps_imag = num_at(field(s, "eps_imag"), 0);
      p.fourier_res[0] = num_at(field(s, "fourier_res1"), 0);
      p.fourier_res[1] = num_at(field(s, "fourier_res2"), 0);
      p.fourier_res[2] = num_at(field(s, "fourier_res3"), 0);
      for (int j = 0; j < 6; ++j) p.ROI[j] = (uint32_t)num_at(field(s, "ROI"), j);
      p.Bornmax = (uint32_t)num_at(field(s, "Bornmax"), 0);
      p.dV = num_at(field(s, "dV"), 0);
      const mxArray* sh = mxGetField(s, 0, "boundary_sharpness");
      p.boundary_sharpness = sh ? num_at(sh, 0) : 1.0;
      const mxArray* ac = mxGetField(s, 0, "acyclic");
      p.acyclic = ac ? (uint32_t)num_at(ac, 0) : 1u;
      const mxArray* sa = mxGetField(s, 0, "shift_axes");
      p.shift_axes = sa ? (uint32_t)num_at(sa, 0) : 7u;
      const mxArray* ra = mxGetField(s, 0, "ramp_axes");
      p.ramp_axes = ra ? (uint32_t)num_at(ra, 0) : 7u;
      const size_t nv = (size_t)e->geo.simulation_size[0] * e->geo.simulation_size[1] * e->geo.simulation_size[2];
      check(cbs_set_potential(e->h, &p, complex_single(prhs[3], nv, "V"), CBS_MEM_HOST));
      break;
    }
    case 2: {  // Born loop from 3D incident fields
      if (nrhs < 3) mexErrMsgIdAndTxt("cbs:arg", "Need to give the field to simulate");
      Entry* e = entry(prhs[1]);
      const size_t nri = (size_t)e->geo.RI_size[0] * e->geo.RI_size[1] * e->geo.RI_size[2] * e->geo.pole_num;
      const size_t n = mxGetNumberOfElements(prhs[2]);
      if (n % nri) mexErrMsgIdAndTxt("cbs:arg", "field size is not a multiple of the RI grid (times the components)");
      const mxComplexSingle* u = complex_single(prhs[2], 0, "u_in");
      plhs[0] = mxCreateUninitNumericArray(mxGetNumberOfDimensions(prhs[2]), mxGetDimensions(prhs[2]),
                                           mxSINGLE_CLASS, mxCOMPLEX);
      check(cbs_solve_3d(e->h, u, (uint32_t)(n / nri), mxGetComplexSingles(plhs[0]), CBS_MEM_HOST));
      break;
    }
    case 3: {  // destroy
      const uint32_t i = (uint32_t)num_at(prhs[1], 0);
      if (i >= g_handles.size() || !g_handles[i]) mexErrMsgIdAndTxt("cbs:id", "Id not found");
      check(cbs_destroy(g_handles[i]->h));
      delete g_handles[i];
      g_handles[i] = nullptr;
      break;
    }
    case 4: {  // 2D in / 2D out
      if (nrhs < 3) mexErrMsgIdAndTxt("cbs:arg", "Need to give the field to simulate");
      Entry* e = entry(prhs[1]);
      const mwSize P = (mwSize)e->geo.pole_num;   // components per angle: [x y (P) A]
      const size_t nf = (size_t)e->geo.field_size[0] * e->geo.field_size[1] * P;
      const size_t n = mxGetNumberOfElements(prhs[2]);
      if (n % nf) mexErrMsgIdAndTxt("cbs:arg", "source size is not a multiple of the field grid (times the components)");
      const mwSize A = (mwSize)(n / nf);
      const mwSize d2[4] = {e->geo.field_size[0], e->geo.field_size[1], P, A};
      const mwSize d3[5] = {e->geo.RI_size[0], e->geo.RI_size[1], e->geo.RI_size[2], P, A};
      plhs[0] = mxCreateUninitNumericArray(4, d2, mxSINGLE_CLASS, mxCOMPLEX);
      void* tr = mxGetComplexSingles(plhs[0]);
      void* rf = nullptr;
      void* f3 = nullptr;
      if (nlhs > 1) {
        plhs[1] = mxCreateUninitNumericArray(4, d2, mxSINGLE_CLASS, mxCOMPLEX);
        rf = mxGetComplexSingles(plhs[1]);
      }
      if (nlhs > 2) {
        plhs[2] = mxCreateUninitNumericArray(5, d3, mxSINGLE_CLASS, mxCOMPLEX);
        f3 = mxGetComplexSingles(plhs[2]);
      }
      check(cbs_solve_2d(e->h, complex_single(prhs[2], 0, "source_2d"), (uint32_t)A, tr, rf, f3, CBS_MEM_HOST));
      break;
    }
    default:
      mexErrMsgIdAndTxt("cbs:arg", "unknown action");
  }
}
