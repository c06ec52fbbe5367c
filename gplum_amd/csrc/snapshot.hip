// snapshot.hip -- snapshots in the reference's ASCII format and the energy-log line.
//
// Header (FileHeader::writeAscii, src/energy.h:89-116): time, n_body, id_next, then
// etot, ekin, ephi_sun, ephi_planet, edisp of the initial and of the current energy.
// One line per particle (FPGrav::writeAscii, src/particle.h:843-855): id, mass, r_planet,
// f, pos, vel, neighbor.number, flag word.  Same printf formats, so files written here
// read back in the reference and in its analysis scripts, and the reference's files
// (e.g. sample/INIT3000.dat) read here.  Host code: the device system hands its
// particles over with get_particles.
#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
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

// FileHeader / FPGrav object layouts of the reference's default build (x86-64)
static_assert(sizeof(gplum_energy) == 56, "Energy is 7 doubles");
static_assert(sizeof(gplum_file_header) == 128 && offsetof(gplum_file_header, time) == 8 &&
                  offsetof(gplum_file_header, e_init) == 16 && offsetof(gplum_file_header, e_now) == 72,
              "FileHeader layout");
static_assert(offsetof(gplum_fpgrav, pos) == 8 && offsetof(gplum_fpgrav, r_out) == 32 &&
                  offsetof(gplum_fpgrav, id) == 48 && offsetof(gplum_fpgrav, vel) == 64 &&
                  offsetof(gplum_fpgrav, acc_d) == 88 && offsetof(gplum_fpgrav, acc) == 112 &&
                  offsetof(gplum_fpgrav, acc_gd) == 208 && offsetof(gplum_fpgrav, phi) == 232 &&
                  offsetof(gplum_fpgrav, v_disp) == 256 && offsetof(gplum_fpgrav, r_out_inv) == 264 &&
                  offsetof(gplum_fpgrav, time) == 272 && offsetof(gplum_fpgrav, r_planet) == 296 &&
                  offsetof(gplum_fpgrav, neighbor) == 312 && offsetof(gplum_fpgrav, id_cluster) == 328 &&
                  offsetof(gplum_fpgrav, n_cluster) == 336 && offsetof(gplum_fpgrav, inDomain) == 340 &&
                  sizeof(gplum_fpgrav) == 344,
              "FPGrav layout");

void to_fpgrav(const gplum_particle &q, int index, gplum_fpgrav &r) {
    std::memset(&r, 0, sizeof(r));
    r.id_local = index;
    r.myrank = 0;
    r.r_out = q.r_out;
    r.r_search = q.r_search;
    r.id = q.id;
    r.mass = q.mass;
    for (int k = 0; k < 3; k++) {
        r.pos[k] = q.pos[k]; r.vel[k] = q.vel[k]; r.acc_d[k] = q.acc_d[k]; r.acc[k] = q.acc[k];
        r.acc_s[k] = q.acc_s[k]; r.jerk_d[k] = q.jerk_d[k]; r.jerk_s[k] = q.jerk_s[k];
    }
    r.phi = q.phi; r.phi_d = q.phi_d; r.phi_s = q.phi_s;
    r.v_disp = q.v_disp; r.r_out_inv = q.r_out_inv;
    r.time = q.time; r.dt = q.dt; r.acc0 = q.acc0;
    r.r_planet = q.r_planet; r.f = q.f;
    r.neighbor = q.neighbor;
    r.id_cluster = q.id_cluster;
    r.n_cluster = q.n_cluster;
    r.inDomain = 1;
    r.isSent = 0;
    r.isDead = (q.flags & GPLUM_FLAG_DEAD) ? 1 : 0;
    r.isMerged = (q.flags & GPLUM_FLAG_MERGED) ? 1 : 0;
}

void from_fpgrav(const gplum_fpgrav &r, gplum_particle &q) {
    std::memset(&q, 0, sizeof(q));
    q.id = r.id;
    q.id_cluster = r.id_cluster;
    q.mass = r.mass; q.r_planet = r.r_planet; q.f = r.f;
    for (int k = 0; k < 3; k++) {
        q.pos[k] = r.pos[k]; q.vel[k] = r.vel[k]; q.acc[k] = r.acc[k]; q.acc_d[k] = r.acc_d[k];
        q.jerk_d[k] = r.jerk_d[k]; q.acc_s[k] = r.acc_s[k]; q.jerk_s[k] = r.jerk_s[k];
    }
    q.phi = r.phi; q.phi_d = r.phi_d; q.phi_s = r.phi_s;
    q.r_out = r.r_out; q.r_out_inv = r.r_out_inv; q.r_search = r.r_search; q.v_disp = r.v_disp;
    q.time = r.time; q.dt = r.dt; q.acc0 = r.acc0;
    q.neighbor = r.neighbor;
    q.n_cluster = r.n_cluster;
    q.flags = (r.isDead ? GPLUM_FLAG_DEAD : 0) | (r.isMerged ? GPLUM_FLAG_MERGED : 0);
}

gplum_energy energy7(const double e[5], double edisp) {
    // e = {etot, ekin, ephi_sun, ephi, ephi_d} (System::energy)
    return gplum_energy{e[0], e[1], e[2], e[3] + e[4], e[3], e[4], edisp};
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
    if ((rc = S->merge_resolve())) return rc;   // energy() has synchronised the stream
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

int gplum_restart_write(const char *path, const gplum_file_header *h, const gplum_particle *pp, int n) {
    if (!path || !h || n < 0 || (n > 0 && !pp)) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    gplum::FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return gplum::fail(GPLUM_E_ARG, std::string("cannot open ") + path);
    gplum_file_header hh = *h;
    hh.n_body = n;
    if (std::fwrite(&hh, sizeof(hh), 1, fc.f) != 1) return gplum::fail(GPLUM_E_ARG, "the header could not be written");
    std::vector<gplum_fpgrav> buf(4096);
    for (int i0 = 0; i0 < n; i0 += (int)buf.size()) {
        const int m = std::min<int>((int)buf.size(), n - i0);
        for (int a = 0; a < m; a++) gplum::to_fpgrav(pp[i0 + a], i0 + a, buf[a]);
        if (std::fwrite(buf.data(), sizeof(gplum_fpgrav), (size_t)m, fc.f) != (size_t)m)
            return gplum::fail(GPLUM_E_ARG, "particle records could not be written");
    }
    if (std::fclose(fc.f) != 0) {
        fc.f = nullptr;
        return gplum::fail(GPLUM_E_ARG, std::string("error closing ") + path);
    }
    fc.f = nullptr;
    return GPLUM_OK;
}

int gplum_restart_read(const char *path, gplum_file_header *h, gplum_particle *pp, int cap, int *n) {
    if (!path || !h || !n) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    gplum::FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return gplum::fail(GPLUM_E_ARG, std::string("cannot open ") + path);
    if (std::fread(h, sizeof(*h), 1, fc.f) != 1 || h->n_body < 0)
        return gplum::fail(GPLUM_E_ARG, "the header could not be read");
    *n = h->n_body;
    if (!pp) return GPLUM_OK;
    if (cap < h->n_body) return gplum::fail(GPLUM_E_OVERFLOW, "particle array smaller than n_body");
    std::vector<gplum_fpgrav> buf(4096);
    for (int i0 = 0; i0 < h->n_body; i0 += (int)buf.size()) {
        const int m = std::min<int>((int)buf.size(), h->n_body - i0);
        if (std::fread(buf.data(), sizeof(gplum_fpgrav), (size_t)m, fc.f) != (size_t)m)
            return gplum::fail(GPLUM_E_ARG, "particle records " + std::to_string(i0) + ".. could not be read");
        for (int a = 0; a < m; a++) gplum::from_fpgrav(buf[a], pp[i0 + a]);
    }
    return GPLUM_OK;
}

int gplum_system_restart_write(gplum_system *s, const char *path, int32_t id_next) {
    if (!s || !path) return gplum::fail(GPLUM_E_ARG, "null argument");
    System *S = reinterpret_cast<System *>(s);
    std::vector<gplum_particle> h(S->n);
    int rc = S->get_particles(h.data(), S->n);
    if (rc) return rc;
    double e[5];
    if ((rc = S->energy(e))) return rc;
    if ((rc = S->merge_resolve())) return rc;   // energy() has synchronised the stream
    gplum_file_header hd;
    std::memset(&hd, 0, sizeof(hd));
    hd.n_body = S->n;
    hd.id_next = id_next;
    hd.time = S->time_sys;
    hd.e_init = gplum::energy7(S->e_init, 0.);
    hd.e_now = gplum::energy7(e, S->edisp);
    return gplum_restart_write(path, &hd, h.data(), S->n);
}

int gplum_system_energy_line(gplum_system *s, char *buf, int cap) {
    if (!s || !buf || cap <= 0) return gplum::fail(GPLUM_E_ARG, "bad arguments");
    System *S = reinterpret_cast<System *>(s);
    double e[5];
    int rc = S->energy(e);
    if (!rc) rc = S->merge_resolve();   // energy() has synchronised the stream
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
