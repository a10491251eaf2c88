// Test harness: read one catalogue with the host layer's ascii reader and
// print every parsed field (hex floats), one object per line.
//   host_reader <param-file> <0|1|2: plain, inverse-p, inverse-p with dither weights> [count-only]
#include <cstdio>
#include <cstdlib>

#include "../two_pcf_amd/csrc/host/twopcf_ref_api.hpp"

int main(int argc, char **argv) {
	if (argc != 3 && argc != 4) return 2;
	Params p(argv[1]);
	const int mode = std::atoi(argv[2]);
	twopcf_enable_inverse_p(mode != 0);
	twopcf_enable_dither_weights(mode == 2);
	GalaxyCatalogue *c = read_file(p.data_filename, p.data_file_type, p);
	std::printf("n %d %d\n", c->n_gals, c->n_words);
	for (int i = 0; i < (argc == 4 ? 0 : c->n_gals); ++i) {
		const Galaxy &g = c->gals[i];
		std::printf("%a %a %a %a %a %a %a %a %d %a", g.x[0], g.x[1], g.x[2], g.x_norm[0], g.x_norm[1], g.x_norm[2],
		            (double)g.weight, (double)g.input_weight, g.jk_region, g.d_comov);
		for (int k = 0; k < c->n_words; ++k) std::printf(" %llu", (unsigned long long)c->masks[(size_t)i * c->n_words + k]);
		if (!c->dither.empty())
			for (int k = 0; k < c->n_words; ++k)
				std::printf(" %llu", (unsigned long long)c->dither[(size_t)i * c->n_words + k]);
		std::printf("\n");
	}
	delete c;
	return 0;
}
