#include "h5lib.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <string>

namespace h5 {

namespace {

struct Loaded {
	Lib api{};
	bool ok = false;
	std::string why;
};

template <class F>
bool sym(void *so, const char *name, F &out, std::string &why) {
	out = reinterpret_cast<F>(dlsym(so, name));
	if (!out) why = std::string("symbol ") + name + " missing";
	return out != nullptr;
}

Loaded load() {
	Loaded L;
	const char *env = std::getenv("TWOPCF_HDF5_LIB");
	const char *names[] = {env, "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
	void *so = nullptr;
	for (const char *n : names)
		if (n && *n && (so = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
	if (!so) {
		L.why = "no libhdf5 found ($TWOPCF_HDF5_LIB, libhdf5.so, libhdf5.so.103, /opt/conda/lib/libhdf5.so)";
		return L;
	}
	Lib &a = L.api;
	std::string &w = L.why;
	bool ok = sym(so, "H5open", a.open, w) && sym(so, "H5Fopen", a.Fopen, w) && sym(so, "H5Fcreate", a.Fcreate, w) &&
	          sym(so, "H5Fclose", a.Fclose, w) && sym(so, "H5Gopen2", a.Gopen2, w) &&
	          sym(so, "H5Gcreate2", a.Gcreate2, w) && sym(so, "H5Gclose", a.Gclose, w) &&
	          sym(so, "H5Dopen2", a.Dopen2, w) && sym(so, "H5Dcreate2", a.Dcreate2, w) &&
	          sym(so, "H5Dread", a.Dread, w) && sym(so, "H5Dwrite", a.Dwrite, w) &&
	          sym(so, "H5Dget_space", a.Dget_space, w) && sym(so, "H5Dclose", a.Dclose, w) &&
	          sym(so, "H5Sget_simple_extent_npoints", a.Sget_simple_extent_npoints, w) &&
	          sym(so, "H5Screate", a.Screate, w) && sym(so, "H5Screate_simple", a.Screate_simple, w) &&
	          sym(so, "H5Sclose", a.Sclose, w) && sym(so, "H5Acreate2", a.Acreate2, w) &&
	          sym(so, "H5Aopen", a.Aopen, w) && sym(so, "H5Awrite", a.Awrite, w) && sym(so, "H5Aread", a.Aread, w) &&
	          sym(so, "H5Aclose", a.Aclose, w) && sym(so, "H5Tcopy", a.Tcopy, w) &&
	          sym(so, "H5Tset_order", a.Tset_order, w) && sym(so, "H5Tclose", a.Tclose, w);
	if (!ok) return L;
	if (a.open() < 0) {
		w = "H5open failed";
		return L;
	}
	// 1.10 keeps the predefined type ids in hid_t globals named <macro>_g
	const char *gn[] = {"H5T_NATIVE_DOUBLE_g", "H5T_NATIVE_FLOAT_g", "H5T_NATIVE_INT_g", "H5T_NATIVE_LONG_g",
	                    "H5T_C_S1_g"};
	hid_t *gv[] = {&a.NATIVE_DOUBLE, &a.NATIVE_FLOAT, &a.NATIVE_INT, &a.NATIVE_LONG, &a.C_S1};
	for (int k = 0; k < 5; ++k) {
		const hid_t *g = reinterpret_cast<const hid_t *>(dlsym(so, gn[k]));
		if (!g) {
			w = std::string("symbol ") + gn[k] + " missing";
			return L;
		}
		*gv[k] = *g;
	}
	L.ok = true;
	return L;
}

const Loaded &loaded() {
	static const Loaded L = load();
	return L;
}

}  // namespace

const Lib &lib() {
	const Loaded &L = loaded();
	if (!L.ok) {
		std::fprintf(stderr, "twopcf: the hdf5 file type needs the HDF5 C library: %s\n", L.why.c_str());
		std::exit(1);
	}
	return L.api;
}

bool available() { return loaded().ok; }

}  // namespace h5
