// C++ host classes over the C ABI; see hero_host.h.
#include "hero_host.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>

namespace hero_host {

Config::Config() { hero_default_options(&opt); }

Device::Device(const Config &cfg) : opt_(cfg.opt) {
	// without -c the reference writes alignment records (.aln); keep them
	opt_.records = opt_.call_consensus ? 0 : 1;
	const int rc = hero_create(&opt_, &ctx_);
	if (rc != HERO_OK) throw std::runtime_error("hero_create failed (no usable GPU or invalid options)");
}

Device::~Device() { hero_destroy(ctx_); }

void Device::check(int rc, const char *what) const {
	if (rc != HERO_OK) throw std::runtime_error(std::string(what) + ": " + hero_last_error(ctx_));
}

void ReadBatch::add(const std::string &s, int64_t uid) {
	offset.push_back((int64_t) seq.size());
	length.push_back((int32_t) s.size());
	unique_id.push_back(uid);
	seq += s;
}

bool ReadBatch::loadFasta(const std::string &path, ReadBatch &out) {
	std::ifstream in(path);
	if (!in) return false;
	std::string name, line;
	while (std::getline(in, name)) {
		if (name.empty()) continue;
		if (name[0] != '>' || !std::getline(in, line)) return false;
		out.add(line, std::atoll(name.c_str() + 1));
	}
	return true;
}

QueryDataset::QueryDataset(Device &dev, const ReadBatch &reads) : dev_(dev), reads_(reads) {
	for (int32_t l : reads_.length) longestReadLength = std::max<uint64_t>(longestReadLength, (uint64_t) l);
	dev_.check(hero_set_queries(dev_.ctx(), reads_.size(), reads_.seq.data(), reads_.offset.data(), reads_.length.data(),
					   reads_.unique_id.data()),
			"hero_set_queries");
}

bool QueryDataset::qualityFilter(const std::string &s) {
	uint64_t cnt[4] = {0, 0, 0, 0};
	for (char c : s) {
		if (c != 'A' && c != 'C' && c != 'G' && c != 'T') return false;
		cnt[((unsigned char) c >> 1) & 3]++;
	}
	const uint64_t thr = (uint64_t) ((double) s.size() * .8);
	return !(cnt[0] >= thr || cnt[1] >= thr || cnt[2] >= thr || cnt[3] >= thr);
}

const std::string *QueryDataset::getReadFromID(int64_t id) {
	if (id < 0 || id >= reads_.size()) {
		std::printf("ID %lld out of bound.\n", (long long) id);
		return nullptr;
	}
	cache_.resize(1);
	cache_[0] = reads_.read(id);
	return &cache_[0];
}

void QueryDataset::callConsensus() {
	std::string out(reads_.seq.size(), '\0');
	dev_.check(hero_call_consensus(dev_.ctx(), &out[0]), "hero_call_consensus");
	reads_.seq.swap(out);
}

std::string QueryDataset::queryOutText() const {
	std::string t;
	t.reserve(reads_.seq.size() + (size_t) reads_.size() * 12);
	for (int64_t i = 0; i < reads_.size(); i++) {
		t += '>';
		t += std::to_string((long long) reads_.unique_id[(size_t) i]);
		t += '\n';
		t.append(reads_.seq, (size_t) reads_.offset[(size_t) i], (size_t) reads_.length[(size_t) i]);
		t += '\n';
	}
	return t;
}

void QueryDataset::writeQueryOut(const std::string &filename) {
	std::ofstream f(filename, std::ios_base::app | std::ios_base::binary);
	f << queryOutText();
}

std::string QueryDataset::coverageText() const {
	// the histogram grows until it holds the largest coverage (hero_coverage_histogram reports it)
	std::vector<uint64_t> counts(1 << 12);
	int32_t maxc = 0;
	for (;;) {
		const int rc = hero_coverage_histogram(dev_.ctx(), counts.data(), (int32_t) counts.size(), &maxc);
		if (rc == HERO_OK) break;
		if (rc != HERO_ERR_LIMIT || maxc >= (1 << 30)) dev_.check(rc, "hero_coverage_histogram");
		counts.assign((size_t) maxc + 1, 0);
	}
	std::string out = "COV\tFRE\n";
	for (int32_t v = 0; v <= maxc; v++)
		if (counts[(size_t) v]) out += std::to_string(v) + "\t" + std::to_string((uint32_t) counts[(size_t) v]) + "\n";
	return out;
}

void QueryDataset::getCoverageAndWriteResults(const std::string &filename) {
	std::ofstream f(filename, std::ios_base::app | std::ios_base::binary);
	f << coverageText();
}

HashTableLongKey::HashTableLongKey(QueryDataset *q) : dataSet(q), hashKeyLength(q->device().options().k), dev_(q->device()) {}

bool HashTableLongKey::insertQueryDataset(QueryDataset *q) {
	dataSet = q;
	dev_.check(hero_build_index(dev_.ctx()), "hero_build_index");
	return true;
}

std::vector<int64_t> HashTableLongKey::getListOfReads(const std::string &s, int64_t start, int64_t len) {
	std::vector<int64_t> out;
	if (len != hashKeyLength || start < 0 || start + len > (int64_t) s.size()) return out;
	int64_t count = 0, off = 0;
	dev_.check(hero_lookup(dev_.ctx(), 1, s.data() + start, &count, &off, nullptr, 0), "hero_lookup");
	if (count <= 0) return out;
	std::vector<uint64_t> e((size_t) count);
	dev_.check(hero_lookup(dev_.ctx(), 1, s.data() + start, &count, &off, e.data(), count), "hero_lookup");
	out.assign(e.begin(), e.end());
	return out;
}

uint64_t HashTableLongKey::hashFunction(const std::string &s, uint64_t P) {
	uint64_t h = 0;
	dev_.check(hero_hash_kmers(dev_.ctx(), 1, (int32_t) s.size(), s.data(), P, &h), "hero_hash_kmers");
	return h;
}

void BandAligner::setSequences(const std::string &S, int64_t s0, int64_t s1, const std::string &Q, int64_t q0,
		int64_t q1) {
	if (s0 < 0 || q0 < 0) {
		std::printf("Invalid start position.\n");
		std::exit(1);
	}
	s_ = S.substr((size_t) s0, (size_t) (s1 - s0));
	q_ = Q.substr((size_t) q0, (size_t) (q1 - q0));
	flag_ = 0; // setSequences resets the left / right flags
}

void BandAligner::calculateAlignment() {
	BandAlignerBatch b(dev_);
	b.add(s_, 0, (int64_t) s_.size(), q_, 0, (int64_t) q_.size(), m_band, flag_);
	std::vector<BandPairResult> r = b.run();
	m_band = r[0].band;
	vCigarCount.swap(r[0].counts);
	vCigarType.swap(r[0].types);
}

void BandAlignerBatch::add(const std::string &S, int64_t s0, int64_t s1, const std::string &Q, int64_t q0, int64_t q1,
		int64_t band, int flag) {
	hero_band_pair p;
	p.s_off = (int64_t) seq_.size() + s0;
	p.s_len = (int32_t) (s1 - s0);
	seq_ += S;
	p.q_off = (int64_t) seq_.size() + q0;
	p.q_len = (int32_t) (q1 - q0);
	seq_ += Q;
	p.band = (int32_t) band;
	p.flag = flag;
	pairs_.push_back(p);
}

std::vector<BandPairResult> BandAlignerBatch::run() {
	const int32_t n = (int32_t) pairs_.size();
	std::vector<int32_t> band(n), cn(n);
	std::vector<int64_t> coff(n);
	int64_t cap = 0;
	for (const hero_band_pair &p : pairs_) cap += (int64_t) p.s_len + p.q_len + 1;
	std::vector<int32_t> counts((size_t) cap);
	std::vector<char> types((size_t) cap);
	dev_.check(hero_band_align_batch(dev_.ctx(), n, pairs_.data(), seq_.data(), band.data(), coff.data(), cn.data(),
					   counts.data(), types.data(), cap),
			"hero_band_align_batch");
	std::vector<BandPairResult> r((size_t) n);
	for (int32_t i = 0; i < n; i++) {
		r[(size_t) i].band = band[(size_t) i];
		r[(size_t) i].counts.assign(counts.begin() + coff[(size_t) i], counts.begin() + coff[(size_t) i] + cn[(size_t) i]);
		r[(size_t) i].types.assign(types.begin() + coff[(size_t) i], types.begin() + coff[(size_t) i] + cn[(size_t) i]);
	}
	return r;
}

bool SubjectDataset::loadNextChunkParallel(uint16_t) {
	chunk = ReadBatch();
	if (next_ >= all_.size()) return false;
	const int64_t end = std::min<int64_t>(all_.size(), next_ + chunk_size_);
	for (int64_t i = next_; i < end; i++) chunk.add(all_.read(i), all_.unique_id[(size_t) i]);
	iCountSubjectRead = (uint64_t) chunk.size();
	next_ = end;
	return true;
}

static void take_alignments(Device &dev, std::string &aln) {
	int64_t n = 0;
	dev.check(hero_take_alignments(dev.ctx(), nullptr, 0, &n), "hero_take_alignments");
	if (n == 0) return;
	const size_t at = aln.size();
	aln.resize(at + (size_t) n);
	dev.check(hero_take_alignments(dev.ctx(), &aln[at], n, &n), "hero_take_alignments");
}

bool SubjectReadAligner::start(SubjectDataset &subjects) {
	Device &dev = query_->device();
	while (subjects.loadNextChunkParallel()) {
		const ReadBatch &c = subjects.chunk;
		dev.check(hero_align_subjects(dev.ctx(), c.size(), c.seq.data(), c.offset.data(), c.length.data(),
						  c.unique_id.data()),
				"hero_align_subjects");
		if (!dev.options().call_consensus) take_alignments(dev, aln_);
	}
	return true;
}

bool SubjectReadAligner::startSelf() {
	Device &dev = query_->device();
	dev.check(hero_align_self(dev.ctx()), "hero_align_self");
	if (!dev.options().call_consensus) take_alignments(dev, aln_);
	return true;
}

} // namespace hero_host
