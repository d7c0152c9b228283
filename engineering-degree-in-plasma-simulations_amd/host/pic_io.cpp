// pic_io.cpp -- the reference's ASCII output formats (Outputs.cpp:9-205)
// written from plain host arrays; see include/pic_io.h.  Numbers go through
// std::ostream with its default formatting, as the reference's do, so the
// same doubles produce the same text.
#include "../../include/pic_io.h"

#include <fstream>
#include <iomanip>
#include <ostream>
#include <string>

namespace {

// Field<T> operator<< (Field.h:676-693): k-planes, j-rows, i fastest, each
// value followed by a space, a newline after every k-plane
template <class T>
void write_scalars(std::ostream& out, const T* a, int ni, int nj, int nk) {
    for (int k = 0; k < nk; k++) {
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) out << a[((size_t)k * nj + j) * ni + i] << " ";
        out << "\n";
    }
}

// a Vec3 (Vec3.h:380-384): three width-5 fields separated by spaces
void write_vec3(std::ostream& out, const double* v) {
    out << std::setw(5) << v[0] << " " << std::setw(5) << v[1] << " " << std::setw(5) << v[2];
}

void write_vectors(std::ostream& out, const double* a, int ni, int nj, int nk) {
    for (int k = 0; k < nk; k++) {
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                write_vec3(out, a + 3 * (((size_t)k * nj + j) * ni + i));
                out << " ";
            }
        out << "\n";
    }
}

void array_head(std::ostream& out, const std::string& name, int comps) {
    out << "<DataArray Name=\"" << name << "\" NumberOfComponents=\"" << comps
        << "\" format=\"ascii\" type=\"Float64\">\n";
}

}  // namespace

extern "C" int pic_io_write_vti(const char* path, const pic_io_fields* f) {
    if (!path || !f) return -1;
    std::ofstream out(path);
    if (!out.is_open()) return -1;
    const int ni = f->ni, nj = f->nj, nk = f->nk;
    out << "<VTKFile type=\"ImageData\">\n";
    out << "<ImageData Origin=\"" << f->x0[0] << " " << f->x0[1] << " " << f->x0[2] << "\" ";
    out << "Spacing=\"" << f->dx[0] << " " << f->dx[1] << " " << f->dx[2] << "\" ";
    out << "WholeExtent=\"0 " << ni - 1 << " 0 " << nj - 1 << " 0 " << nk - 1 << "\">\n";
    out << "<PointData>\n";
    array_head(out, "NodeVol", 1);
    write_scalars(out, f->node_vol, ni, nj, nk);
    out << "</DataArray>\n";
    array_head(out, "ObjectID", 1);
    write_scalars(out, f->object_id, ni, nj, nk);
    out << "</DataArray>\n";
    array_head(out, "NodeType", 1);
    write_scalars(out, f->node_type, ni, nj, nk);
    out << "</DataArray>\n";
    array_head(out, "phi", 1);
    write_scalars(out, f->phi, ni, nj, nk);
    out << "</DataArray>\n";
    array_head(out, "rho", 1);
    write_scalars(out, f->rho, ni, nj, nk);
    out << "</DataArray>\n";
    const pic_io_species* sp = f->species;
    for (int s = 0; s < f->nspecies; s++) {
        array_head(out, std::string("nd.") + sp[s].name, 1);
        write_scalars(out, sp[s].den, ni, nj, nk);
        out << "</DataArray>\n";
    }
    for (int s = 0; s < f->nspecies; s++) {
        array_head(out, std::string("avg_nd.") + sp[s].name, 1);
        write_scalars(out, sp[s].den_avg, ni, nj, nk);
        out << "</DataArray>\n";
    }
    for (int s = 0; s < f->nspecies; s++) {
        array_head(out, std::string("vel.") + sp[s].name, 3);
        write_vectors(out, sp[s].vel, ni, nj, nk);
        out << "</DataArray>\n";
    }
    for (int s = 0; s < f->nspecies; s++) {
        array_head(out, std::string("T.") + sp[s].name, 1);
        write_scalars(out, sp[s].T, ni, nj, nk);
        out << "</DataArray>\n";
    }
    array_head(out, "ef", 3);
    write_vectors(out, f->ef, ni, nj, nk);
    out << "</DataArray>\n";
    out << "</PointData>\n";
    out << "<CellData>\n";
    for (int s = 0; s < f->nspecies; s++) {
        array_head(out, std::string("mpc.") + sp[s].name, 1);
        write_scalars(out, sp[s].mpc, ni - 1, nj - 1, nk - 1);
        out << "</DataArray>\n";
    }
    out << "</CellData>\n";
    out << "</ImageData>\n";
    out << "</VTKFile>\n";
    return out.good() ? 0 : -1;
}

extern "C" int pic_io_write_vtp(const char* path, const char* name, const double* aos7, size_t n,
                                int count) {
    if (!path || !name || (n && !aos7)) return -1;
    std::ofstream out(path);
    if (!out.is_open()) return -1;
    // Outputs.cpp:231-256: every step-th particle, step = n / count
    size_t take = 0, step = 1;
    if (n != 0 && count > 0) {
        step = n / (size_t)count;
        take = (size_t)count;
        if ((size_t)count > n) {
            take = n;
            step = 1;
        }
    }
    size_t m = 0;
    for (size_t k = 0; k < take; k++)
        if (k * step < n) m++;
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "<PolyData>\n";
    out << "<Piece NumberOfPoints=\"" << m << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" ";
    out << "NumberOfStrips=\"0\" NumberOfCells=\"0\">\n";
    out << "<Points>\n";
    out << "<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (size_t k = 0; k < take; k++)
        if (k * step < n) {
            write_vec3(out, aos7 + 7 * (k * step));
            out << "\n";
        }
    out << "</DataArray>\n";
    out << "</Points>\n";
    out << "<PointData>\n";
    out << "<DataArray Name=\"vel." << name << "\" type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (size_t k = 0; k < take; k++)
        if (k * step < n) {
            write_vec3(out, aos7 + 7 * (k * step) + 3);
            out << "\n";
        }
    out << "</DataArray>\n";
    out << "</PointData>\n";
    out << "</Piece>\n";
    out << "</PolyData>\n";
    out << "</VTKFile>\n";
    return out.good() ? 0 : -1;
}

extern "C" int pic_io_diag_row(const char* path, int header, int nspecies, const char* const* names,
                               int ts, double time, double wall_time, const double* diag, double PE) {
    if (!path || (nspecies > 0 && (!names || !diag))) return -1;
    std::ofstream out(path, header ? std::ios::out | std::ios::trunc : std::ios::out | std::ios::app);
    if (!out.is_open()) return -1;
    if (header) {
        out << "ts,time,wall_time";
        for (int s = 0; s < nspecies; s++) {
            const std::string n = names[s];
            out << ",mp_count." << n << ",real_count." << n << ",px." << n << ",py." << n << ",pz." << n
                << ",KE." << n;
        }
        out << ",PE,E_total" << std::endl;
    }
    out << ts << "," << time;
    out << "," << wall_time;
    double tot_KE = 0;
    for (int s = 0; s < nspecies; s++) {
        const double* d = diag + 6 * s;
        tot_KE += d[5];
        // the reference prints getNumParticles() (a size_t) and the sums
        out << "," << (unsigned long long)d[0] << "," << d[1] << "," << d[2] << "," << d[3] << "," << d[4]
            << "," << d[5];
    }
    out << "," << PE << "," << (tot_KE + PE);
    out << "\n";
    return out.good() ? 0 : -1;
}
