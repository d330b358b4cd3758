/*
 * SWIG interface of the MI355X cross-power library.  Module name, inline
 * helpers and pointer/array helpers follow sphradiometer's
 * src/python/sphradiometer.i (lines 20-138) so Python code written against
 * the reference module keeps working; the LAL time-series helpers of the
 * reference are not part of this path.  Build where swig is available:
 *
 *   swig -python -I../include sphradiometer.i
 *   cc -shared -fPIC sphradiometer_wrap.c -I../include $(python3-config --includes) \
 *      -L. -lsphradiometer_amd -o _sphradiometer.so
 *
 * Where swig is absent, sphradiometer_amd/sphradiometer.py provides the
 * same names over ctypes.
 */
%module sphradiometer
%include carrays.i
%include ccomplex.i
%include cpointer.i

%{
#define SWIG_FILE_WITH_INIT
#include <sphradiometer/instrument.h>
#include <sphradiometer/sh_series.h>
#include <sphradiometer/projection.h>
#include <sphradiometer/correlator.h>
#include <sphradiometer/diagnostics.h>
#include <sphradiometer/inspiral_sky_map.h>
#include <sphradiometer/sphradiometer_gpu.h>
#include <sphradiometer/sphradiometer_batch.h>
%}

%inline %{
	/* element i of the plans array of a network plan (sphradiometer.i:45-47) */
	struct correlator_plan_fd *pick_ith_correlator_plan_fd(struct correlator_plan_fd **array, int i) {
		return array[i];
	}
%}

%include <sphradiometer/sphr_compat.h>
%include <sphradiometer/instrument.h>
%include <sphradiometer/sh_series.h>
%include <sphradiometer/projection.h>
%include <sphradiometer/correlator.h>
%include <sphradiometer/diagnostics.h>
%include <sphradiometer/inspiral_sky_map.h>
%include <sphradiometer/sphradiometer_gpu.h>
%include <sphradiometer/sphradiometer_batch.h>

%pointer_functions(unsigned int, uintp);
%pointer_functions(struct correlator_network_plan_fd, correlator_network_plan_fdp);
%pointer_functions(struct autocorrelator_network_plan_fd, autocorrelator_network_plan_fdp);
%pointer_functions(struct sh_series, sh_seriesp);
%pointer_functions(struct sh_series *, sh_seriespp);

%array_functions(double, double_array);
%array_functions(double *, doublep_array);
%array_functions(double complex, double_complex_array);
%array_functions(double complex *, double_complexp_array);
