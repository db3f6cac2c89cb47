// A C++ caller written against the reference's engine interface (src/AbstractHawkes.hpp:32-64,
// the calls of src/benchmark.cpp:154-211): factory, setTimesData, updateLocations,
// setParameters, getSumOfLikContribs, getProbsSelfExcite, plus getLogLikeGradient of the
// extended interface hph::HawkesEngine (reached from the factory's pointer by a cast). Reads
// "D N" then N rows of D locations and a time, then 6 parameters from stdin; prints the
// log-likelihood, the gradient and the probabilities with 17 significant digits.
#include "hph_engine.hpp"

#include <cstdio>
#include <vector>

int main() {
    int D = 0, N = 0;
    if (std::scanf("%d %d", &D, &N) != 2 || D < 1 || N < 1) return 2;
    std::vector<double> locations((size_t)N * D), times(N), params(6);
    for (int i = 0; i < N; ++i) {
        for (int d = 0; d < D; ++d)
            if (std::scanf("%lf", &locations[(size_t)i * D + d]) != 1) return 2;
        if (std::scanf("%lf", &times[i]) != 1) return 2;
    }
    for (double &p : params)
        if (std::scanf("%lf", &p) != 1) return 2;
    try {
        hph::SharedPtr engine = hph::factory(D, N, hph::OPENCL, 0, 1);
        engine->setTimesData(times.data(), times.size());
        engine->updateLocations(-1, locations.data(), locations.size());
        engine->setParameters(params.data(), params.size());
        const double ll = engine->getSumOfLikContribs();
        std::vector<double> grad(6), probs(N);
        auto ext = std::dynamic_pointer_cast<hph::HawkesEngine>(engine);
        if (!ext) {
            std::fprintf(stderr, "error: the factory's engine does not implement hph::HawkesEngine\n");
            return 1;
        }
        ext->getLogLikeGradient(grad.data(), grad.size());
        engine->getProbsSelfExcite(probs.data(), probs.size());
        std::printf("%.17g\n", ll);
        for (double g : grad) std::printf("%.17g ", g);
        std::printf("\n");
        for (double p : probs) std::printf("%.17g\n", p);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
