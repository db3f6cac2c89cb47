// An engine subclass written against the reference's base class (src/AbstractHawkes.hpp:32-60):
// the nine pure virtuals and the protected members embeddingDimension, locationCount,
// observationCount, flags and updatedLocation. It must compile against include/hph_engine.hpp
// unchanged. Prints the members of an instance at N = 100,000, where the reference's int
// observationCount would have overflowed.
#include "hph_engine.hpp"

#include <cstdio>

class CountingEngine : public hph::AbstractHawkes {
  public:
    CountingEngine(int D, int N, long flags) : hph::AbstractHawkes(D, N, flags) {}

    double getSumOfLikContribs() override { return (double)observationCount; }
    void updateLocations(int index, double *, size_t) override { updatedLocation = index; }
    void getProbsSelfExcite(double *out, size_t length) override {
        for (size_t i = 0; i < length; ++i) out[i] = 0.0;
    }
    void storeState() override {}
    void restoreState() override {}
    void acceptState() override { updatedLocation = -1; }
    void setTimesData(double *, size_t) override {}
    void setParameters(double *, size_t) override {}
    int getInternalDimension() override { return embeddingDimension; }

    long flagsValue() const { return flags; }
    int count() const { return locationCount; }
    int updated() const { return updatedLocation; }
};

int main() {
    hph::SharedPtr e = std::make_shared<CountingEngine>(2, 100000, hph::OPENCL);
    double x[2] = {0.0, 0.0};
    e->updateLocations(7, x, 2);
    auto *c = static_cast<CountingEngine *>(e.get());
    std::printf("%d %d %.0f %ld %d\n", e->getInternalDimension(), c->count(), e->getSumOfLikContribs(),
                c->flagsValue(), c->updated());
    e->acceptState();
    return c->updated() == -1 ? 0 : 1;
}
