// Drives a disk through the C++ host interface (fargocpt_amd/csrc/fcpt_host.hpp)
// the way the reference's main() and sim::run drive theirs, for the GPU
// parity test tests/test_gpu_host_api.py.
//
//   host_api_check <dir> <nsteps> <stepwise|device>
// <dir> holds desc.bin (fcpt_disk_desc bytes, radii pointer ignored),
// radii.bin, sigma.bin, vrad.bin, vazi.bin, energy.bin and bodies.bin (per
// step: n_planets x 4 doubles, then the indirect term); results are written
// back as out_sigma.bin, out_vrad.bin, out_vazi.bin, out_energy.bin and
// out_time.bin (time, last_dt).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../fargocpt_amd/csrc/fcpt_host.hpp"

static std::vector<char> slurp(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path.c_str());
        std::exit(2);
    }
    std::vector<char> buf;
    char tmp[65536];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    return buf;
}

static std::vector<double> doubles(const std::string &path)
{
    std::vector<char> b = slurp(path);
    std::vector<double> v(b.size() / sizeof(double));
    std::memcpy(v.data(), b.data(), v.size() * sizeof(double));
    return v;
}

static void dump(const std::string &path, const std::vector<double> &v)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc < 4) return 2;
    const std::string dir = argv[1];
    const unsigned long nsteps = std::strtoul(argv[2], nullptr, 10);
    const bool stepwise = std::string(argv[3]) == "stepwise";
    std::vector<char> db = slurp(dir + "/desc.bin");
    if (db.size() != sizeof(fcpt_disk_desc)) {
        std::fprintf(stderr, "descriptor size %zu != %zu\n", db.size(), sizeof(fcpt_disk_desc));
        return 2;
    }
    fcpt_disk_desc desc;
    std::memcpy(&desc, db.data(), sizeof(desc));
    std::vector<double> radii = doubles(dir + "/radii.bin");
    desc.radii = radii.data();
    std::vector<double> bodies = doubles(dir + "/bodies.bin");
    const int nb = (int)(bodies.size() / nsteps - 2) / 4;
    try {
        fcpt::Data data(desc);
        const auto sigma = doubles(dir + "/sigma.bin"), vrad = doubles(dir + "/vrad.bin"),
                   vazi = doubles(dir + "/vazi.bin"), energy = doubles(dir + "/energy.bin");
        data.upload_reference(sigma, vrad, vazi, energy);
        data.upload(sigma, vrad, vazi, energy);
        // main(): first CalculateTimeStep (main.cpp:117), then sim::init
        fcpt::sim::time = 0.0;
        fcpt::sim::last_dt = 0.1;  // FirstDT
        fcpt::sim::N_monitor = 0;
        fcpt::sim::CalculateTimeStep(data);
        fcpt::sim::init(data);
        unsigned long step = 0;
        auto hook = [&](fcpt::Data &d, double) {
            const double *rec = bodies.data() + step * (4 * nb + 2);
            std::vector<fcpt::Body> b(nb);
            for (int k = 0; k < nb; ++k) b[k] = {rec[4 * k], rec[4 * k + 1], rec[4 * k + 2], rec[4 * k + 3]};
            d.set_bodies(b, rec[4 * nb], rec[4 * nb + 1]);
            ++step;
        };
        if (stepwise) fcpt::sim::run_stepwise(data, nsteps, hook);
        else fcpt::sim::run(data, nsteps, hook);
        dump(dir + "/out_sigma.bin", data.field(FCPT_SIGMA));
        dump(dir + "/out_vrad.bin", data.field(FCPT_VRAD));
        dump(dir + "/out_vazi.bin", data.field(FCPT_VAZI));
        dump(dir + "/out_energy.bin", data.field(FCPT_ENERGY));
        dump(dir + "/out_time.bin", {fcpt::sim::time, fcpt::sim::last_dt});
    } catch (const fcpt::Error &e) {
        std::fprintf(stderr, "fcpt error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
