// Link check of the reference-side adapter (gpu_hydro.cpp): a program that
// takes the address of every adapter function, so that the link resolves
// each reference symbol and each fcpt_* symbol they use.  It is built, not run.
#include <cstdint>
#include <cstdio>
#include <string>

class t_data;
namespace gpu_hydro
{
void init(t_data &data, int32_t precision);
void init_on_device();
void publish_bodies(t_data &data, double t);
void download(t_data &data);
double calculate_time_step(t_data &data);
void step_euler_gas_before_transport(t_data &data, double dt);
void step_euler_gas_after_transport(double dt);
double run_iteration();
void compute_disk_on_nbody_accel(t_data &data);
void write_snapshot(uint32_t fields);
void read_snapshot(const std::string &dir, const std::string &reference_dir);
void finish();
}  // namespace gpu_hydro

int main()
{
    const void *fns[] = {
        (const void *)&gpu_hydro::init, (const void *)&gpu_hydro::init_on_device,
        (const void *)&gpu_hydro::publish_bodies, (const void *)&gpu_hydro::download,
        (const void *)&gpu_hydro::calculate_time_step,
        (const void *)&gpu_hydro::step_euler_gas_before_transport,
        (const void *)&gpu_hydro::step_euler_gas_after_transport,
        (const void *)&gpu_hydro::run_iteration, (const void *)&gpu_hydro::compute_disk_on_nbody_accel,
        (const void *)&gpu_hydro::write_snapshot, (const void *)&gpu_hydro::read_snapshot,
        (const void *)&gpu_hydro::finish};
    int n = 0;
    for (const void *p : fns) n += p != nullptr;
    std::printf("%d adapter functions linked\n", n);
    return 0;
}
