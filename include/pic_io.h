/*
 * pic_io.h -- output writers of the PIC library's host side (no GPU).
 *
 * The reference's Output namespace (ch4/v3/src/Outputs.cpp) writes ASCII
 * VTK ImageData of the node fields, VTK PolyData of sampled particles and a
 * runtime diagnostics CSV.  These functions write the same files -- element
 * names, array names, value order and iostream default number formatting --
 * from plain host arrays, so ParaView states and CSV tooling built for the
 * reference read them unchanged.  The C++ host classes' Output::
 * functions fill the arrays from the device and call these.
 *
 * Conventions: node arrays i-fastest (u = (k nj + j) ni + i), vector fields
 * xyz-interleaved per node, cell arrays c = (k (nj-1) + j) (ni-1) + i,
 * particles 7 doubles {x, y, z, vx, vy, vz, w}.  Return 0, or -1 when the
 * file cannot be written.
 */
#ifndef PIC_IO_H
#define PIC_IO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const char* name;
    const double* den;      /* number density (Species::den) */
    const double* den_avg;  /* moving average (Species::den_avg) */
    const double* vel;      /* stream velocity, 3 per node (Species::vel) */
    const double* T;        /* temperature (Species::T) */
    const double* mpc;      /* macroparticles per cell (Species::macro_part_count) */
} pic_io_species;

typedef struct {
    int ni, nj, nk;
    double x0[3], dx[3];
    const double* node_vol;
    const int32_t* object_id;
    const int32_t* node_type;
    const double* phi;
    const double* rho;
    const double* ef;  /* 3 per node */
    int nspecies;
    const pic_io_species* species;
} pic_io_fields;

/* Output::fieldsOutput (Outputs.cpp:9-123): the ImageData file */
int pic_io_write_vti(const char* path, const pic_io_fields* f);

/* Output::particlesOutput (Outputs.cpp:207-321) for one species: particles
 * k * (n / count) for k < count (all n when count > n), positions as Points
 * and velocities as the PointData array vel.<name> */
int pic_io_write_vtp(const char* path, const char* name, const double* aos7, size_t n, int count);

/* Output::diagOutput (Outputs.cpp:143-179): with header != 0 the file is
 * created with the column header (ts,time,wall_time, per species
 * mp_count,real_count,px,py,pz,KE, then PE,E_total), then one row is
 * appended.  diag holds per species {mp_count, real_count, px, py, pz, KE}. */
int pic_io_diag_row(const char* path, int header, int nspecies, const char* const* names, int ts,
                    double time, double wall_time, const double* diag, double PE);

#ifdef __cplusplus
}
#endif
#endif
