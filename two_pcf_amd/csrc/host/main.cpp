// twopcf / twopcf_invp — the parameter-file driver (main.cpp:12-75 of the
// reference) on the GPU counting library.  Usage: twopcf <param-file>
#include <cstdlib>
#include <iostream>
#include <string>

#include "twopcf_ref_api.hpp"

// The reference selects the inverse-p pipeline when it is compiled
// (_USE_INV_WEIGHTS, makefile:4); the twopcf_invp driver is this file built
// with TWOPCF_INVERSE_P=1.
#ifndef TWOPCF_INVERSE_P
#define TWOPCF_INVERSE_P 0
#endif

int main(int argc, char *argv[]) {
	std::cout << "--TWOPCF correlation function calculator (MI355X)--\n\n";
	if (argc != 2) {
		std::cout << "Incorrect number of arguments given, exiting\n";
		return 1;
	}
	std::cout << "Correct number of arguments passed\n";
	Params params(argv[1]);
	twopcf_enable_inverse_p(TWOPCF_INVERSE_P);
	const char *pi = std::getenv("TWOPCF_PI");
	if (pi && std::string(pi) == "midpoint") twopcf_set_pi_definition(TWOPCF_PI_MIDPOINT);
	// dither weights (_USE_DITHER_WEIGHTS, makefile:6) in the inverse-p pipeline
	const char *dither = std::getenv("TWOPCF_DITHER");
	if (TWOPCF_INVERSE_P && dither && std::string(dither) == "1") twopcf_enable_dither_weights(true);
	GalaxyCatalogue *data = read_file(params.data_filename, params.data_file_type, params);
	GalaxyCatalogue *randoms = read_file(params.random_filename, params.random_file_type, params);
	if (TWOPCF_INVERSE_P) {
		// the inverse-p pipeline (main.cpp:49-62 under _USE_INV_WEIGHTS)
		calc_angular_upweight(data, randoms, params);
		if (params.plot_monopole) auto_correlate_monopole_invpweights(data, randoms, params);
		if (params.plot_sigma_pi) auto_correlate_sigma_pi_invpweights(data, randoms, params);
		if (params.plot_s_mu) auto_correlate_s_mu_invpweights(data, randoms, params);
	} else {
		if (params.plot_monopole) auto_correlate_monopole(data, randoms, params);
		if (params.plot_sigma_pi) auto_correlate_sigma_pi(data, randoms, params);
		if (params.plot_s_mu) auto_correlate_s_mu(data, randoms, params);
	}
	delete data;
	delete randoms;
	return 0;
}
