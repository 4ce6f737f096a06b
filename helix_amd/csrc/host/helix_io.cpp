// helix_io.cpp -- energy-trace and checkpoint files (include/helix_io.h, include/helix/io.h).
#include "helix_io.h"

#include <helix/io.h>
#include <helix/types.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <ostream>
#include <string>
#include <vector>

namespace {

thread_local std::string g_io_error;

int io_fail(const std::string& msg)
{
	g_io_error = msg;
	return -1;
}

constexpr char kMagic[8] = {'H', 'L', 'X', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kHeaderBytes = 96;

uint64_t fnv1a(const void* data, size_t bytes, uint64_t h = 1469598103934665603ull)
{
	const unsigned char* p = static_cast<const unsigned char*>(data);
	for(size_t q = 0; q < bytes; ++q) {
		h ^= p[q];
		h *= 1099511628211ull;
	}
	return h;
}

template <typename X>
void put(std::vector<unsigned char>& b, size_t off, X v)
{
	std::memcpy(b.data() + off, &v, sizeof(X));
}

template <typename X>
X get(const unsigned char* b, size_t off)
{
	X v;
	std::memcpy(&v, b + off, sizeof(X));
	return v;
}

// one "<time> <energy>" row as the legacy CLI prints it (std::setprecision(8), default float format)
void energy_row(std::ostream& os, double time, double energy, bool float32)
{
	os << std::setprecision(8) << time << " ";
	if(float32)
		os << static_cast<float>(energy);
	else
		os << energy;
	os << "\n";
}

}  // namespace

extern "C" {

const char* helix_io_last_error(void)
{
	return g_io_error.c_str();
}

int helix_io_write_energy_trace(const char* path, const double* times, const double* energies, int64_t rows,
	int float32_values)
{
	if(!path || rows < 0 || (rows > 0 && (!times || !energies))) return io_fail("bad arguments");
	std::ofstream os(path);
	if(!os) return io_fail(std::string("cannot open ") + path);
	for(int64_t r = 0; r < rows; ++r) energy_row(os, times[r], energies[r], float32_values != 0);
	os.flush();
	if(!os) return io_fail(std::string("write failed: ") + path);
	return 0;
}

uint64_t helix_io_hierarchy_hash(const int32_t* rows, int64_t n_ado, int32_t n_modes)
{
	if(!rows || n_ado <= 0 || n_modes <= 0) return 0;
	return fnv1a(rows, static_cast<size_t>(n_ado) * static_cast<size_t>(n_modes) * sizeof(int32_t));
}

uint64_t helix_io_problem_hash(const helix_hip_problem* pr)
{
	if(!pr || pr->n <= 0 || pr->n_ado <= 0 || pr->n_modes <= 0 || pr->n_baths <= 0) return 0;
	const size_t n = static_cast<size_t>(pr->n), H = static_cast<size_t>(pr->n_ado);
	const size_t M = static_cast<size_t>(pr->n_modes), B = static_cast<size_t>(pr->n_baths);
	uint64_t h = 1469598103934665603ull;
	auto add = [&h](const void* p, size_t bytes) {
		if(p && bytes) h = fnv1a(p, bytes, h);
	};
	const int32_t dims[5] = {pr->n, pr->n_ado, pr->n_modes, pr->n_baths, pr->order};
	add(dims, sizeof(dims));
	add(&pr->dt, sizeof(double));
	add(pr->mode_bath, M * sizeof(int32_t));
	add(pr->down_comm, H * M * 2 * sizeof(double));
	add(pr->down_anti, H * M * 2 * sizeof(double));
	add(pr->self_coef, H * 2 * sizeof(double));
	add(pr->xi, B * 2 * sizeof(double));
	add(pr->xi_counter, B * 2 * sizeof(double));
	auto add_spin = [&](const helix_hip_spin_operator& o) {
		add(o.diag, n * sizeof(double));
		add(&o.nflips, sizeof(int32_t));
		if(o.nflips > 0) {
			add(o.flip_masks, static_cast<size_t>(o.nflips) * sizeof(int32_t));
			add(o.flip_coefs, static_cast<size_t>(o.nflips) * sizeof(double));
		}
	};
	auto add_csr = [&](const helix_hip_csr_operator& o) {
		add(&o.nnz, sizeof(int32_t));
		add(o.row_offsets, (n + 1) * sizeof(int32_t));
		if(o.nnz > 0) {
			add(o.columns, static_cast<size_t>(o.nnz) * sizeof(int32_t));
			add(o.values, static_cast<size_t>(o.nnz) * 2 * sizeof(double));
		}
	};
	add_spin(pr->hamiltonian);
	add(pr->hamiltonian_dense, n * n * 2 * sizeof(double));
	if(pr->hamiltonian_csr) add_csr(*pr->hamiltonian_csr);
	for(size_t b = 0; b < B; ++b) {
		if(pr->couplings) add_spin(pr->couplings[b]);
		if(pr->couplings_csr) add_csr(pr->couplings_csr[b]);
	}
	return h;
}

int helix_io_checkpoint_write(const char* path, const helix_io_checkpoint* info, const double* state)
{
	if(!path || !info || !state) return io_fail("bad arguments");
	if(info->n <= 0 || info->n_ado <= 0) return io_fail("empty state");
	const uint64_t payload = static_cast<uint64_t>(info->n_ado) * static_cast<uint64_t>(info->n) *
		static_cast<uint64_t>(info->n) * 16ull;
	std::vector<unsigned char> h(kHeaderBytes, 0);
	std::memcpy(h.data(), kMagic, 8);
	put<uint32_t>(h, 8, kVersion);
	put<uint32_t>(h, 12, kHeaderBytes);
	const int32_t dims[6] = {info->n, info->n_ado, info->n_modes, info->n_baths, info->order, info->precision};
	std::memcpy(h.data() + 16, dims, sizeof(dims));
	put<uint64_t>(h, 40, info->steps);
	put<double>(h, 48, info->time);
	put<double>(h, 56, info->dt);
	put<uint64_t>(h, 64, info->hierarchy_hash);
	put<uint64_t>(h, 72, fnv1a(state, payload));
	put<uint64_t>(h, 80, payload);
	put<uint64_t>(h, 88, info->problem_hash);
	const std::string tmp = std::string(path) + ".tmp";
	{
		std::ofstream os(tmp, std::ios::binary);
		if(!os) return io_fail("cannot open " + tmp);
		os.write(reinterpret_cast<const char*>(h.data()), kHeaderBytes);
		os.write(reinterpret_cast<const char*>(state), static_cast<std::streamsize>(payload));
		os.flush();
		if(!os) return io_fail("write failed: " + tmp);
	}
	// a complete file replaces the old one in one step
	if(std::rename(tmp.c_str(), path) != 0) return io_fail(std::string("rename failed: ") + path);
	return 0;
}

int helix_io_checkpoint_read(const char* path, helix_io_checkpoint* info, double* state, uint64_t state_doubles)
{
	if(!path || !info) return io_fail("bad arguments");
	std::ifstream is(path, std::ios::binary);
	if(!is) return io_fail(std::string("cannot open ") + path);
	unsigned char h[kHeaderBytes];
	is.read(reinterpret_cast<char*>(h), kHeaderBytes);
	if(is.gcount() != static_cast<std::streamsize>(kHeaderBytes)) return io_fail("truncated header");
	if(std::memcmp(h, kMagic, 8) != 0) return io_fail("not a helix checkpoint (magic)");
	if(get<uint32_t>(h, 8) != kVersion) return io_fail("unsupported checkpoint version");
	if(get<uint32_t>(h, 12) != kHeaderBytes) return io_fail("unexpected header size");
	int32_t dims[6];
	std::memcpy(dims, h + 16, sizeof(dims));
	info->n = dims[0];
	info->n_ado = dims[1];
	info->n_modes = dims[2];
	info->n_baths = dims[3];
	info->order = dims[4];
	info->precision = dims[5];
	info->steps = get<uint64_t>(h, 40);
	info->time = get<double>(h, 48);
	info->dt = get<double>(h, 56);
	info->hierarchy_hash = get<uint64_t>(h, 64);
	info->problem_hash = get<uint64_t>(h, 88);
	const uint64_t hash = get<uint64_t>(h, 72);
	const uint64_t payload = get<uint64_t>(h, 80);
	if(info->n <= 0 || info->n_ado <= 0 ||
		payload != static_cast<uint64_t>(info->n_ado) * static_cast<uint64_t>(info->n) * static_cast<uint64_t>(info->n) * 16ull)
		return io_fail("payload size does not match the header");
	if(!state) return 0;
	if(state_doubles * 8ull < payload) return io_fail("state buffer too small");
	is.read(reinterpret_cast<char*>(state), static_cast<std::streamsize>(payload));
	if(static_cast<uint64_t>(is.gcount()) != payload) return io_fail("truncated payload");
	if(fnv1a(state, payload) != hash) return io_fail("payload hash mismatch (corrupted checkpoint)");
	return 0;
}

}  // extern "C"

namespace helix {

void write_energy_trace(std::ostream& out, const RunResult& result, EnergyValues values)
{
	const std::size_t rows = std::min(result.times.size(), result.energy_diagonal.size());
	for(std::size_t r = 0; r < rows; ++r)
		energy_row(out, result.times[r], result.energy_diagonal[r], values == EnergyValues::Float32);
}

bool write_energy_trace(const std::string& path, const RunResult& result, EnergyValues values)
{
	std::ofstream os(path);
	if(!os) return false;
	write_energy_trace(os, result, values);
	os.flush();
	return static_cast<bool>(os);
}

}  // namespace helix
