// HDF5 C library (1.10 ABI), opened at run time.
//
// The reference links libhdf5 when it is built with _HAVE_HDF5 (makefile:3)
// and then reads catalogues (io.cpp:417-590), writes correlation functions
// (io.cpp:653-975) and hands angular histograms between runs
// (hist.cpp:115-221) through it.  The host layer here opens the library only
// when a parameter file asks for the hdf5 file type, so the drivers start
// without it and no other shared library is pulled from its directory.
//
// Search order: $TWOPCF_HDF5_LIB, then libhdf5.so / libhdf5.so.103 on the
// loader's path, then /opt/conda/lib/libhdf5.so.
#pragma once

#include <cstdint>

namespace h5 {

// H5Ipublic.h, H5public.h (1.10)
typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

// H5Fpublic.h, H5Ppublic.h, H5Spublic.h, H5Tpublic.h (1.10)
constexpr unsigned ACC_RDONLY = 0x0000u;
constexpr unsigned ACC_TRUNC = 0x0002u;
constexpr hid_t P_DEFAULT = 0;
constexpr hid_t S_ALL = 0;
constexpr int S_SCALAR = 0;   // H5S_class_t
constexpr int T_ORDER_LE = 0;  // H5T_order_t

struct Lib {
	herr_t (*open)();
	hid_t (*Fopen)(const char *, unsigned, hid_t);
	hid_t (*Fcreate)(const char *, unsigned, hid_t, hid_t);
	herr_t (*Fclose)(hid_t);
	hid_t (*Gopen2)(hid_t, const char *, hid_t);
	hid_t (*Gcreate2)(hid_t, const char *, hid_t, hid_t, hid_t);
	herr_t (*Gclose)(hid_t);
	hid_t (*Dopen2)(hid_t, const char *, hid_t);
	hid_t (*Dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
	herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
	herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
	hid_t (*Dget_space)(hid_t);
	herr_t (*Dclose)(hid_t);
	int64_t (*Sget_simple_extent_npoints)(hid_t);
	hid_t (*Screate)(int);
	hid_t (*Screate_simple)(int, const hsize_t *, const hsize_t *);
	herr_t (*Sclose)(hid_t);
	hid_t (*Acreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
	hid_t (*Aopen)(hid_t, const char *, hid_t);
	herr_t (*Awrite)(hid_t, hid_t, const void *);
	herr_t (*Aread)(hid_t, hid_t, void *);
	herr_t (*Aclose)(hid_t);
	hid_t (*Tcopy)(hid_t);
	herr_t (*Tset_order)(hid_t, int);
	herr_t (*Tclose)(hid_t);
	// predefined types (the H5T_*_g globals, valid after H5open)
	hid_t NATIVE_DOUBLE, NATIVE_FLOAT, NATIVE_INT, NATIVE_LONG, C_S1;
};

// The loaded library; exits with a message naming the search order when no
// libhdf5 can be opened.
const Lib &lib();

// Whether a libhdf5 can be opened (no exit).
bool available();

}  // namespace h5
