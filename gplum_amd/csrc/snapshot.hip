// snapshot.hip -- snapshots in the reference's ASCII format and the energy-log line.
//
// Header (FileHeader::writeAscii, src/energy.h:89-116): time, n_body, id_next, then
// etot, ekin, ephi_sun, ephi_planet, edisp of the initial and of the current energy.
// One line per particle (FPGrav::writeAscii, src/particle.h:843-855): id, mass, r_planet,
// f, pos, vel, neighbor.number, flag word.  Same printf formats, so files written here
// read back in the reference and in its analysis scripts, and the reference's files
// (e.g. sample/INIT3000.dat) read here.  Host code: the device system hands its
// particles over with get_particles.
#include <cstdio>
#include <cstring>
#include <vector>

#include "system.hpp"

namespace gplum {
namespace {

struct FileCloser {
    FILE *f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

int write_header(FILE *fp, const gplum_snapshot_header &h) {
    const int r = std::fprintf(fp,
                               "%g\t%d\t%d\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t"
                               "%20.15e\t%20.15e\n",
                               h.time, h.n_body, h.id_next, h.e_init[0], h.e_init[1], h.e_init[2], h.e_init[3],
                               h.e_init[4], h.e_now[0], h.e_now[1], h.e_now[2], h.e_now[3], h.e_now[4]);
    return r > 0 ? GPLUM_OK : fail(GPLUM_E_ARG, "the snapshot header could not be written");
}

int write_particle(FILE *fp, const gplum_particle &q) {
    const int r = std::fprintf(fp,
                               "%lld\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t%20.15e\t"
                               "%d\t%d\n",
                               (long long)q.id, q.mass, q.r_planet, q.f, q.pos[0], q.pos[1], q.pos[2], q.vel[0],
                               q.vel[1], q.vel[2], (int)q.neighbor.number, 0);
    return r > 0 ? GPLUM_OK : fail(GPLUM_E_ARG, "a snapshot particle line could not be written");
}

int read_header(FILE *fp, gplum_snapshot_header &h) {
    std::memset(&h, 0, sizeof(h));
    const int r = std::fscanf(fp, "%lf\t%d\t%d\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\n", &h.time, &h.n_body,
                              &h.id_next, &h.e_init[0], &h.e_init[1], &h.e_init[2], &h.e_init[3], &h.e_init[4],
                              &h.e_now[0], &h.e_now[1], &h.e_now[2], &h.e_now[3], &h.e_now[4]);
    if (r != 13 || h.n_body < 0) return fail(GPLUM_E_ARG, "the snapshot header could not be read");
    return GPLUM_OK;
}

}  // namespace
}  // namespace gplum

using gplum::System;

extern "C" {

int gplum_snapshot_write(const char *path, const gplum_snapshot_header *h, const gplum_particle *pp, int n) {
    if (!path || !h || n < 0 || (n > 0 && !pp)) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    gplum::FileCloser fc{std::fopen(path, "w")};
    if (!fc.f) return gplum::fail(GPLUM_E_ARG, std::string("cannot open ") + path);
    gplum_snapshot_header hh = *h;
    hh.n_body = n;
    int rc = gplum::write_header(fc.f, hh);
    for (int i = 0; i < n && !rc; i++) rc = gplum::write_particle(fc.f, pp[i]);
    if (rc) return rc;
    if (std::fclose(fc.f) != 0) {
        fc.f = nullptr;
        return gplum::fail(GPLUM_E_ARG, std::string("error closing ") + path);
    }
    fc.f = nullptr;
    return GPLUM_OK;
}

int gplum_snapshot_read(const char *path, gplum_snapshot_header *h, gplum_particle *pp, int cap, int *n) {
    if (!path || !h || !n) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    gplum::FileCloser fc{std::fopen(path, "r")};
    if (!fc.f) return gplum::fail(GPLUM_E_ARG, std::string("cannot open ") + path);
    int rc = gplum::read_header(fc.f, *h);
    if (rc) return rc;
    *n = h->n_body;
    if (!pp) return GPLUM_OK;   // header only: the caller sizes its array
    if (cap < h->n_body) return gplum::fail(GPLUM_E_OVERFLOW, "particle array smaller than n_body");
    for (int i = 0; i < h->n_body; i++) {
        gplum_particle &q = pp[i];
        std::memset(&q, 0, sizeof(q));
        long long id = 0;
        int number = 0, flag = 0;
        const int r = std::fscanf(fc.f, "%lld\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%d\t%d\n", &id, &q.mass,
                                  &q.r_planet, &q.f, &q.pos[0], &q.pos[1], &q.pos[2], &q.vel[0], &q.vel[1],
                                  &q.vel[2], &number, &flag);
        if (r != 12)
            return gplum::fail(GPLUM_E_ARG, "particle line " + std::to_string(i) + " could not be read");
        q.id = id;
        q.id_cluster = id;
        q.neighbor.number = number;
    }
    return GPLUM_OK;
}

int gplum_system_snapshot(gplum_system *s, const char *path, int32_t id_next) {
    if (!s || !path) return gplum::fail(GPLUM_E_ARG, "null argument");
    System *S = reinterpret_cast<System *>(s);
    std::vector<gplum_particle> h(S->n);
    int rc = S->get_particles(h.data(), S->n);
    if (rc) return rc;
    double e[5];
    if ((rc = S->energy(e))) return rc;
    gplum_snapshot_header hd;
    std::memset(&hd, 0, sizeof(hd));
    hd.time = S->time_sys;
    hd.n_body = S->n;
    hd.id_next = id_next;
    // Energy: etot, ekin, ephi_sun, ephi_planet = ephi + ephi_d, edisp
    hd.e_init[0] = S->e_init[0];
    hd.e_init[1] = S->e_init[1];
    hd.e_init[2] = S->e_init[2];
    hd.e_init[3] = S->e_init[3] + S->e_init[4];
    hd.e_init[4] = 0.;
    hd.e_now[0] = e[0];
    hd.e_now[1] = e[1];
    hd.e_now[2] = e[2];
    hd.e_now[3] = e[3] + e[4];
    hd.e_now[4] = S->edisp;
    return gplum_snapshot_write(path, &hd, h.data(), S->n);
}

int gplum_system_energy_line(gplum_system *s, char *buf, int cap) {
    if (!s || !buf || cap <= 0) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    System *S = reinterpret_cast<System *>(s);
    double e[5];
    int rc = S->energy(e);
    if (rc) return rc;
    // outputStep (src/func.h:87-130): time, n_tot, etot, dE = (etot - etot0 - edisp)/etot0,
    // largest cluster, number of clusters, isolated particles
    const double de = (e[0] - S->e_init[0] - S->edisp) / S->e_init[0];
    const long long n_iso = (long long)S->n - S->stats.n_cluster_particles;
    const int w = std::snprintf(buf, (size_t)cap, "%.8f\t%d\t%.15e\t%.15e\t%lld\t%lld\t%lld\n", S->time_sys, S->n,
                                e[0], de, (long long)S->stats.max_cluster, (long long)S->stats.n_clusters, n_iso);
    if (w < 0 || w >= cap) return gplum::fail(GPLUM_E_OVERFLOW, "energy line longer than the buffer");
    return GPLUM_OK;
}

}  // extern "C"
