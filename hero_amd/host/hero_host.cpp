// C++ host classes over the C ABI; see hero_host.h.
#include "hero_host.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>

#include "../csrc/hd.h"

namespace hero_host {

Config::Config() { hero_default_options(&opt); }

hero_options Config::deviceOptions() const {
	hero_options o = opt;
	o.records = o.call_consensus ? 0 : 1; // without -c the reference writes alignment records (.aln)
	o.long_format = bSpeicialVersion ? 1 : 0;
	o.merge = (bMerging && o.call_consensus) ? 1 : 0;
	return o;
}

Device::Device(const Config &cfg) : cfg_(cfg), opt_(cfg.deviceOptions()) {
	if (opt_.long_format && opt_.call_consensus)
		throw std::runtime_error("-v is supported for alignment records (.aln), not with -c");
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

// ---- files ------------------------------------------------------------------------------------
// One file in memory, read with std::getline's semantics: a read that finds its delimiter stops
// after it; one that reaches the end takes what is left and sets eof; a read at the end yields "".
struct TextFile {
	std::string buf;
	size_t pos = 0;
	bool eof = false;
	void getline(std::string &out, char delim) {
		out.clear();
		if (pos >= buf.size()) {
			eof = true;
			return;
		}
		const size_t q = buf.find(delim, pos);
		if (q == std::string::npos) {
			out.assign(buf, pos, std::string::npos);
			pos = buf.size();
			eof = true;
		} else {
			out.assign(buf, pos, q - pos);
			pos = q + 1;
		}
	}
};

struct ReadFiles::Impl {
	std::vector<std::string> files;
	size_t index = 0;
	TextFile f;
	int type = -1; // 0 FASTA, 1 FASTQ, decided by the first file's first line
	bool open(size_t i) {
		std::ifstream in(files[i], std::ios_base::binary);
		if (!in) return false;
		f = TextFile();
		f.buf.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
		if (type < 0) {
			std::string first;
			TextFile probe = f;
			probe.getline(first, '\n');
			type = (!first.empty() && first[0] == '@') ? 1 : 0; // '>' FASTA; the reference warns on others
		}
		return true;
	}
	// one record: header line, sequence; `text` is the record's last read (the reference drops a
	// record whose last read is empty where its loops end)
	void record(std::string &name, std::string &seq, std::string &text) {
		std::string line0;
		f.getline(line0, '\n');
		if (type == 0) {
			f.getline(seq, '>');
			seq.erase(std::remove(seq.begin(), seq.end(), '\n'), seq.end());
			text = seq; // the reference tests the sequence with its newlines removed
		} else {
			f.getline(seq, '\n');
			std::string plus;
			f.getline(plus, '\n');
			f.getline(text, '\n');
		}
		name = (!line0.empty() && (line0[0] == '>' || line0[0] == '@')) ? line0.substr(1) : line0;
		for (char &c : seq) c = (char) std::toupper((unsigned char) c);
	}
};

ReadFiles::ReadFiles(const std::vector<std::string> &files) : p_(new Impl()) {
	p_->files = files;
	for (const std::string &fn : files) {
		std::ifstream in(fn);
		if (!in) ok_ = false;
	}
	if (files.empty() || !ok_ || !p_->open(0)) ok_ = false, p_->f.eof = true, p_->index = files.size();
}

ReadFiles::~ReadFiles() {}

// The query loaders' loop (QueryDataset.cpp:103-391): per file, records until its end; the file's
// last record is dropped when its last read was empty.
bool ReadFiles::next(std::string &name, std::string &sequence) {
	for (;;) {
		if (p_->index >= p_->files.size()) return false;
		if (!p_->f.eof) {
			std::string text;
			p_->record(name, sequence, text);
			if (!(p_->f.eof && text.empty())) return true;
		}
		// this file is done: the next one
		if (++p_->index >= p_->files.size()) return false;
		if (!p_->open(p_->index)) throw std::runtime_error("Unable to open file: " + p_->files[p_->index]);
	}
}

bool ReadFiles::nextRaw(std::string &name, std::string &sequence, std::string &text) {
	if (p_->index >= p_->files.size()) return false;
	if (p_->f.eof) {
		if (++p_->index >= p_->files.size()) return false;
		if (!p_->open(p_->index)) throw std::runtime_error("Unable to open file: " + p_->files[p_->index]);
	}
	p_->record(name, sequence, text);
	return true;
}

bool ReadBatch::loadFasta(const std::string &path, ReadBatch &out) {
	ReadFiles rf({path});
	if (!rf.ok()) return false;
	std::string name, s;
	while (rf.next(name, s)) out.add(s, std::atoll(name.c_str()));
	return true;
}

// ---- QueryDataset -------------------------------------------------------------------------------
QueryDataset::QueryDataset(Device &dev, const std::string &queryFilename) : dev_(dev), files_{queryFilename} {}

QueryDataset::QueryDataset(Device &dev, const std::vector<std::string> &queryFilenames) : dev_(dev), files_(queryFilenames) {}

QueryDataset::QueryDataset(Device &dev, const ReadBatch &reads) : dev_(dev), reads_(reads) { upload(); }

bool QueryDataset::buildDataset() {
	if (files_.size() == 1) return loadDatasetParallel(files_[0], dev_.config().numberOfThreads);
	return loadDatasetParallel(files_, dev_.config().numberOfThreads);
}

bool QueryDataset::loadDatasetParallel(const std::string &filename, uint16_t threads) {
	return loadDatasetParallel(std::vector<std::string>{filename}, threads);
}

bool QueryDataset::loadDatasetParallel(const std::vector<std::string> &filenames, uint16_t) {
	ReadFiles rf(filenames);
	if (!rf.ok()) {
		std::printf("Unable to open file: %s\n", filenames.empty() ? "" : filenames[0].c_str());
		return false;
	}
	const Config &cfg = dev_.config();
	// uniqueId = atoi(name); under -v (and, reading several files, --mapping) the index, with the
	// name kept (QueryDataset.cpp:211-216,356-361)
	const bool by_index = cfg.bSpeicialVersion || (filenames.size() > 1 && cfg.opt.mapping);
	reads_ = ReadBatch();
	std::string name, s;
	for (int64_t i = 0; rf.next(name, s); i++) {
		reads_.add(s, by_index ? i : std::atoll(name.c_str()));
		if (cfg.bSpeicialVersion) reads_.name.push_back(name);
	}
	upload();
	return true;
}

void QueryDataset::upload() {
	const int64_t n = reads_.size();
	longestReadLength = 0;
	raw_len_.resize((size_t) n);
	good_.resize((size_t) n);
	merged_.assign((size_t) n, 0);
	stat_.assign((size_t) n, Stat());
	const int64_t min_ov = (uint16_t) dev_.options().min_overlap;
	for (int64_t i = 0; i < n; i++) {
		const std::string r = reads_.read(i);
		raw_len_[(size_t) i] = (uint32_t) r.size();
		longestReadLength = std::max<uint64_t>(longestReadLength, r.size());
		good_[(size_t) i] = (int64_t) r.size() > min_ov && qualityFilter(r);
		double gc = 0; // QueryRead::setGcContent (QueryRead.cpp:73-81), of the read as loaded
		for (char c : r) gc += (c == 'C' || c == 'G') ? 1 : 0;
		stat_[(size_t) i].dGcContent = gc / (double) r.size();
	}
	dev_.check(hero_set_queries(dev_.ctx(), n, reads_.seq.data(), reads_.offset.data(), reads_.length.data(),
					   reads_.unique_id.data()),
			"hero_set_queries");
	loaded_ = true;
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
	const std::string before = dev_.config().bCoverageStatistic ? reads_.seq : std::string();
	std::string out(reads_.seq.size(), '\0');
	dev_.check(hero_call_consensus(dev_.ctx(), &out[0]), "hero_call_consensus");
	reads_.seq.swap(out);
	if (!dev_.config().bCoverageStatistic) return;
	// QueryRead::calculateCoverageStatistic (QueryRead.cpp:192-212) of every good read: per base the
	// sum of its four counters, self vote included; min, median (mean of the middle two for an even
	// length, integer division), max. dActualMismatchRate: bases the consensus changed / length.
	std::vector<uint32_t> tally(reads_.seq.size() * 4);
	dev_.check(hero_get_tally(dev_.ctx(), tally.data()), "hero_get_tally");
	std::vector<uint32_t> v;
	for (int64_t i = 0; i < reads_.size(); i++) {
		if (!good_[(size_t) i]) continue;
		const size_t o = (size_t) reads_.offset[(size_t) i], L = (size_t) reads_.length[(size_t) i];
		v.resize(L);
		double changed = 0;
		for (size_t p = 0; p < L; p++) {
			const uint32_t *t = &tally[4 * (o + p)];
			v[p] = t[0] + t[1] + t[2] + t[3] + 1;
			changed += before[o + p] != reads_.seq[o + p] ? 1 : 0;
		}
		std::sort(v.begin(), v.end());
		Stat &st = stat_[(size_t) i];
		st.iMin = v.front(), st.iMax = v.back();
		st.iMed = L % 2 == 0 ? (v[L / 2] + v[L / 2 - 1]) / 2 : v[(L - 1) / 2];
		st.dActualMismatchRate = changed / (double) L;
	}
}

// ---- --merge --------------------------------------------------------------------------------
static char complement_char(char c) { return (c & 0x02) ? (char) (c ^ 0x04) : (char) (c ^ 0x15); }

// QueryRead::getExtendSequence (QueryRead.cpp:383-401): per extension position the first largest
// count (A, C, T, G order), until a position has none.
static std::string extension_of(const uint32_t *ext) {
	static const char kDNA[4] = {'A', 'C', 'T', 'G'};
	std::string s;
	for (int i = 0; i < 20; i++) {
		int id = -1;
		double largest = 0;
		for (int j = 0; j < 4; j++)
			if (largest < (double) ext[4 * i + j]) largest = (double) ext[4 * i + j], id = j;
		if (largest < 1) break;
		s.push_back(kDNA[id]);
	}
	return s;
}

void QueryDataset::mergePairedReads() {
	const int64_t n = reads_.size();
	if (!dev_.options().merge) throw std::runtime_error("mergePairedReads needs --merge (Config::bMerging) with -c");
	std::vector<uint32_t> ext((size_t) n * 80);
	dev_.check(hero_get_extension_tally(dev_.ctx(), ext.data()), "hero_get_extension_tally");
	std::vector<std::string> seq((size_t) n);
	for (int64_t i = 0; i < n; i++) seq[(size_t) i] = reads_.read(i);
	if (n % 2) throw std::out_of_range("mergePairedReads: an odd number of reads (reads 2i and 2i+1 are mates)");
	for (int64_t i = 0; i < n; i += 2) {
		// QueryRead::mergingPairedReads (QueryRead.cpp:268-317) of left = read i, right = read i + 1
		std::string &left = seq[(size_t) i], &right = seq[(size_t) i + 1];
		const std::string sl = extension_of(&ext[(size_t) i * 80]);
		std::string sr = extension_of(&ext[(size_t) (i + 1) * 80]);
		for (char &c : sr) c = complement_char(c);
		const int lenL = (int) raw_len_[(size_t) i], lenR = (int) raw_len_[(size_t) i + 1];
		auto cbase = [&](int p) { return complement_char(right[(size_t) p]); };
		int pl = 0;
		bool match = true;
		for (; pl < lenL; pl++) {
			match = true;
			for (int a = pl, b = lenR - 1; b >= 0 && a < lenL; b--, a++)
				if (left.at((size_t) a) != cbase(b)) {
					match = false;
					break;
				}
			if (match) break;
		}
		if (!(match && pl < lenL)) continue;
		if ((int) (sl.size() + sr.size()) + (lenL - pl) < 40) continue;
		int mism = 0;
		for (int a = lenR - lenL + pl - 1, j = 0; a >= 0 && j < (int) sl.size(); j++, a--)
			if (cbase(a) != sl[(size_t) j]) mism++;
		for (int a = pl - 1, j = 0; a >= 0 && j < (int) sr.size(); j++, a--)
			if (left.at((size_t) a) != sr[(size_t) j]) mism++;
		if (mism > 0) continue;
		for (int a = lenR - lenL + pl - 1; a >= 0; a--) left.push_back(cbase(a));
		merged_[(size_t) i] = merged_[(size_t) i + 1] = 1;
		right.clear();
	}
	// the reads take their merged sequences
	ReadBatch r;
	for (int64_t i = 0; i < n; i++) {
		r.add(seq[(size_t) i], reads_.unique_id[(size_t) i]);
		if (!reads_.name.empty()) r.name.push_back(reads_.name[(size_t) i]);
	}
	reads_.seq.swap(r.seq), reads_.offset.swap(r.offset), reads_.length.swap(r.length);
}

// ---- output files -----------------------------------------------------------------------------
void QueryDataset::writeRecord(std::string &out, int64_t i) const {
	out += '>';
	out += std::to_string((long long) reads_.unique_id[(size_t) i]);
	out += '\n';
	out.append(reads_.seq, (size_t) reads_.offset[(size_t) i], (size_t) reads_.length[(size_t) i]);
	out += '\n';
}

std::string QueryDataset::queryOutText() const {
	std::string t;
	t.reserve(reads_.seq.size() + (size_t) reads_.size() * 12);
	for (int64_t i = 0; i < reads_.size(); i++) writeRecord(t, i);
	return t;
}

// One line of <name>.info (QueryDataset.cpp:494-505): uniqueId, min / median / max coverage, raw
// length, the consensus's changed-base rate and the GC content, tab-separated, doubles as ostream
// writes them.
std::string QueryDataset::infoLine(int64_t i) const {
	const Stat &st = stat_[(size_t) i];
	std::ostringstream o;
	o << reads_.unique_id[(size_t) i] << "\t" << st.iMin << "\t" << st.iMed << "\t" << st.iMax << "\t"
	  << raw_len_[(size_t) i] << "\t" << st.dActualMismatchRate << "\t" << st.dGcContent << "\n";
	return o.str();
}

void QueryDataset::writeQueryOut(const std::string &filename) {
	const Config &cfg = dev_.config();
	const int64_t n = reads_.size();
	// QueryRead::isGoodRead() returns true before its tests (QueryRead.cpp:214-223): every read is written
	if (cfg.bMerging) {
		std::string merged, unmerged;
		if (n % 2) throw std::out_of_range("writeQueryOut: an odd number of reads with --merge");
		for (int64_t i = 0; i < n; i += 2) {
			if (merged_[(size_t) i] && merged_[(size_t) i + 1]) {
				writeRecord(merged, i);
			} else {
				writeRecord(unmerged, i);
				writeRecord(unmerged, i + 1);
			}
		}
		std::ofstream(filename + "_merged.fasta", std::ios_base::app | std::ios_base::binary) << merged;
		std::ofstream(filename + "_unmerged.fasta", std::ios_base::app | std::ios_base::binary) << unmerged;
	} else {
		std::string t;
		if (cfg.bMergedReads) {
			t = queryOutText();
		} else {
			// reads in pairs: the reference reads past the last one when the count is odd
			if (n % 2) throw std::out_of_range("writeQueryOut: an odd number of reads without --singleReads");
			t = queryOutText();
		}
		std::ofstream(filename + ".fasta", std::ios_base::app | std::ios_base::binary) << t;
	}
	if (cfg.bCoverageStatistic) {
		if (!cfg.bMergedReads && n % 2) throw std::out_of_range("writeQueryOut: an odd number of reads in pairs");
		std::string info;
		for (int64_t i = 0; i < n; i++) info += infoLine(i);
		std::ofstream(filename + ".info", std::ios_base::app | std::ios_base::binary) << info;
	}
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

// ---- HashTableLongKey ---------------------------------------------------------------------------
HashTableLongKey::HashTableLongKey(QueryDataset *q) : dev_(q->device()) {
	dataSet = q;
	hashKeyLength = q->device().options().k;
	const Config &cfg = q->device().config();
	iBitOfReadId = (cfg.bSpeicialVersion || cfg.opt.mapping) ? 35 : 40; // HashTableLongKey.cpp:15-20
	iMask = (1ll << iBitOfReadId) - 1;
}

bool HashTableLongKey::createHashTables() {
	if (!dataSet) {
		std::printf("no data set\n");
		return false;
	}
	const hero_options &o = dev_.options();
	uint64_t units = dataSet->getNumberOfReads();
	if (o.mapping) {
		uint64_t bases = 0;
		for (int32_t l : dataSet->reads().length) bases += (uint64_t) l;
		units = bases / (uint64_t) o.k;
	}
	setHashTableSizeAndInitialize((int64_t) getPrimeLargerThanNumber((uint64_t) o.hash_factor * units));
	return true;
}

void HashTableLongKey::setHashTableSizeAndInitialize(int64_t size) {
	hashTableSize = (uint64_t) size;
	lists_.clear();
	occupied_.clear();
	placed_ = false;
}

uint64_t HashTableLongKey::getPrimeLargerThanNumber(uint64_t number) { return hero_prime_larger_than(number); }

bool HashTableLongKey::insertQueryDataset(QueryDataset *q) {
	dataSet = q;
	if (!hashTableSize) createHashTables();
	dev_.check(hero_build_index(dev_.ctx()), "hero_build_index");
	lists_.clear();
	placed_ = false;
	return true;
}

bool HashTableLongKey::insertQueryRead(int64_t, const std::string &, int64_t) {
	std::printf("insertQueryRead: the device index takes whole datasets (insertQueryDataset)\n");
	return false;
}

std::vector<int64_t> *HashTableLongKey::getListOfReads(const std::string &s, int64_t start, int64_t len) {
	if (len != hashKeyLength || start < 0 || start + len > (int64_t) s.size()) return nullptr;
	const std::string key = s.substr((size_t) start, (size_t) len);
	auto it = lists_.find(key);
	if (it != lists_.end()) return it->second.get();
	int64_t count = 0, off = 0;
	dev_.check(hero_lookup(dev_.ctx(), 1, key.data(), &count, &off, nullptr, 0), "hero_lookup");
	if (count <= 0) return nullptr;
	std::vector<uint64_t> e((size_t) count);
	dev_.check(hero_lookup(dev_.ctx(), 1, key.data(), &count, &off, e.data(), count), "hero_lookup");
	std::unique_ptr<std::vector<int64_t>> v(new std::vector<int64_t>());
	for (uint64_t x : e) v->push_back((int64_t) (((x >> 40) << iBitOfReadId) | (x & ((1ull << 40) - 1))));
	std::vector<int64_t> *p = v.get();
	lists_.emplace(key, std::move(v));
	return p;
}

// HashTableLongKey::hashFunction (HashTableLongKey.cpp:168-206): leading-1 sums of the first 32 and
// the remaining bases, ((s1 % P) * (s2 % P)) % P in u64.
static uint64_t host_hash(const char *s, int64_t len, uint64_t P) {
	uint64_t hi = 0, lo = 0;
	for (int64_t i = 0; i < len; i++) {
		const uint64_t c = (uint64_t) hero::ascii_code(s[i]);
		if (i < 32) hi = (hi << 2) | c;
		else lo = (lo << 2) | c;
	}
	return hero::hash_kmer((int) len, hi, lo, P);
}

uint64_t HashTableLongKey::hashFunction(const std::string &sub, uint64_t P) { return host_hash(sub.data(), (int64_t) sub.size(), P); }

uint64_t HashTableLongKey::hashFunction(const std::string &sub) { return hashFunction(sub, hashTableSize); }

uint64_t HashTableLongKey::hashFunction(const std::string &s, int64_t start, int64_t len) {
	return host_hash(s.data() + start, len, hashTableSize);
}

// The reference's insertion (HashTableLongKey.cpp:250-342): every good read in innerId order, keys at
// 0, k, 2k, ... and L - k, each placed by linear probing from its hash at the first slot that is
// empty or holds the same key.
void HashTableLongKey::placeKeys() {
	const uint64_t P = hashTableSize;
	occupied_.assign((size_t) ((P + 63) / 64), 0);
	std::vector<int32_t> owner((size_t) P, -1);
	std::unordered_map<std::string, int32_t> ids;
	numberOfHashCollision = 0, maxSingleHashCollision = 0;
	const ReadBatch &R = dataSet->reads();
	const int k = (int) hashKeyLength;
	for (int64_t r = 0; r < R.size(); r++) {
		if (!dataSet->isGood(r)) continue;
		const std::string s = R.read(r);
		const int n = hero::num_keys((int) s.size(), k);
		for (int t = 0; t < n; t++) {
			const std::string key = s.substr((size_t) hero::key_pos((int) s.size(), k, t), (size_t) k);
			const int32_t id = ids.emplace(key, (int32_t) ids.size()).first->second;
			uint64_t index = host_hash(key.data(), k, P), coll = 0;
			while (owner[(size_t) index] >= 0 && owner[(size_t) index] != id) {
				numberOfHashCollision++, coll++;
				index = index == P - 1 ? 0 : index + 1;
				if (coll >= P) throw std::runtime_error("The size of hash table is too small.");
			}
			owner[(size_t) index] = id;
			occupied_[(size_t) (index >> 6)] |= 1ull << (index & 63);
			maxSingleHashCollision = std::max<uint64_t>(maxSingleHashCollision, coll);
		}
	}
	placed_ = true;
}

bool HashTableLongKey::isEmptyAt(int64_t i) {
	if (i < 0 || (uint64_t) i >= hashTableSize) throw std::out_of_range("isEmptyAt");
	if (!placed_) placeKeys();
	return !((occupied_[(size_t) (i >> 6)] >> (i & 63)) & 1);
}

// ---- BandAligner --------------------------------------------------------------------------------
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

// ---- SubjectDataset -----------------------------------------------------------------------------
SubjectDataset::SubjectDataset(const Config &cfg)
		: files_(new ReadFiles(cfg.vSubjectFiles)), chunk_size_((int64_t) cfg.streamChunkSize),
		  min_overlap_(cfg.opt.min_overlap), keep_names_(cfg.bSpeicialVersion) {
	if (!files_->ok()) throw std::runtime_error("Unable to open a subject file");
}

SubjectDataset::SubjectDataset(const ReadBatch &all, int64_t chunk_size) : all_(&all), chunk_size_(chunk_size) {}

SubjectDataset::~SubjectDataset() {}

bool SubjectDataset::clearSubjectRead() {
	chunk = ReadBatch();
	good.clear();
	iCountSubjectRead = 0;
	return true;
}

bool SubjectDataset::loadNextChunkParallel(uint16_t) {
	clearSubjectRead();
	iLongestReadLength = 0;
	if (all_) {
		if (next_ >= all_->size()) return false;
		const int64_t end = std::min<int64_t>(all_->size(), next_ + chunk_size_);
		for (int64_t i = next_; i < end; i++) {
			const std::string r = all_->read(i);
			chunk.add(r, all_->unique_id[(size_t) i]);
			good.push_back(r.size() < (size_t) min_overlap_ || QueryDataset::qualityFilter(r));
		}
		iCountSubjectRead = (uint64_t) chunk.size();
		next_ = end;
		return true;
	}
	// up to streamChunkSize records, continuing into the next file; the chunk's last record is
	// dropped when its last read was empty (SubjectDataset.cpp:79-121). The reference fails on the
	// call after a chunk that ends exactly at the last file's end; here that call returns false.
	std::string name, s, text;
	std::vector<std::string> names;
	std::vector<std::string> seqs;
	while ((int64_t) seqs.size() < chunk_size_ && files_->nextRaw(name, s, text)) names.push_back(name), seqs.push_back(s);
	if (!seqs.empty() && text.empty()) names.pop_back(), seqs.pop_back();
	for (size_t i = 0; i < seqs.size(); i++) {
		// uniqueId = atoi(name); under -v it stays -1 and the name is kept (SubjectDataset.cpp:150-155)
		chunk.add(seqs[i], keep_names_ ? -1 : std::atoll(names[i].c_str()));
		if (keep_names_) chunk.name.push_back(names[i]);
		good.push_back(seqs[i].size() < (size_t) (uint16_t) min_overlap_ || QueryDataset::qualityFilter(seqs[i]));
		iLongestReadLength = std::max<uint64_t>(iLongestReadLength, seqs[i].size());
	}
	iCountSubjectRead = (uint64_t) chunk.size();
	return chunk.size() > 0;
}

// ---- SubjectReadAligner -------------------------------------------------------------------------
SubjectReadAligner::SubjectReadAligner(HashTableMethod *table) : query_(table->dataSet), table_(table) {}

static void take_alignments(Device &dev, std::string &aln) {
	int64_t n = 0;
	dev.check(hero_take_alignments(dev.ctx(), nullptr, 0, &n), "hero_take_alignments");
	if (n == 0) return;
	const size_t at = aln.size();
	aln.resize(at + (size_t) n);
	dev.check(hero_take_alignments(dev.ctx(), &aln[at], n, &n), "hero_take_alignments");
}

static std::string rc_of(const std::string &s) {
	std::string r(s.rbegin(), s.rend());
	for (char &c : r) c = complement_char(c);
	return r;
}

std::string longFormatRecord(const std::string &queryName, const std::string &query, int strand, int64_t pos,
		const uint32_t *ops_in, int nops, int64_t subjectLength) {
	std::vector<int64_t> cnt;
	std::vector<char> typ;
	for (int i = 0; i < nops; i++) cnt.push_back((int64_t) (ops_in[i] >> 8)), typ.push_back((char) (ops_in[i] & 0xFF));
	// AlignmentRecord::refineAlignment (AlignmentRecord.cpp:318-337): end deletions removed, the
	// position moved past a leading one
	if (!typ.empty() && typ.front() == 'D') pos += cnt.front(), typ.erase(typ.begin()), cnt.erase(cnt.begin());
	if (!typ.empty() && typ.back() == 'D') typ.pop_back(), cnt.pop_back();
	if (typ.empty()) throw std::out_of_range("a record of deletions only");
	int64_t on_ref = 0;
	for (size_t i = 0; i < typ.size(); i++)
		if (typ[i] != 'I') on_ref += cnt[i];
	// getStartPositionOnRef / getEndPositionOnRef / getRelativePosition (:263-284,304-316)
	const int64_t start = typ.front() == 'I' ? pos - cnt.front() : pos;
	const int64_t end = pos + on_ref + (typ.back() == 'I' ? cnt.back() : 0) - 1;
	const char rel = start < 0 ? (end < subjectLength ? 'L' : 'O') : (end < subjectLength ? 'I' : 'R');
	// counts as the reference defines them (:199-231): "deletion" = interior and last I ops less the
	// last one, "insert" = D ops, mismatches = substituted bases
	int64_t del = 0, ins = 0, mis = 0;
	for (size_t i = 1; i < typ.size(); i++)
		if (typ[i] == 'I') del += cnt[i];
	if (typ.back() == 'I') del -= cnt.back();
	for (size_t i = 0; i < typ.size(); i++) {
		if (typ[i] == 'D') ins += cnt[i];
		if (typ[i] != 'D' && typ[i] != 'I' && typ[i] != 'M') mis += cnt[i];
	}
	std::ostringstream o;
	o << queryName << "\t" << rel << "\t" << (del + ins + mis) << "\t" << (strand == 0 ? 'F' : 'R') << "\t"
	  << (typ.front() == 'I' ? cnt.front() : 0) << "\t" << (typ.back() == 'I' ? cnt.back() : 0) << "\t" << query.size()
	  << "\t" << start << "\t" << end << "\t" << mis << "\t" << ins << "\t" << del << "\t";
	// getCigarOnRef (:151-197): I and D exchanged, end insertions left out, then the interior
	// inserted bases
	if (typ.front() != 'I') o << cnt.front() << typ.front();
	const int last = (int) typ.size() - 1;
	for (int i = 1; i < last; i++) o << cnt[(size_t) i] << (typ[(size_t) i] == 'D' ? 'I' : typ[(size_t) i] == 'I' ? 'D' : typ[(size_t) i]);
	if (typ.size() > 1 && typ.back() != 'I') o << cnt.back() << (typ.back() == 'D' ? 'I' : typ.back());
	o << ",";
	size_t index = 0;
	for (size_t i = 0; i < typ.size(); i++) {
		if (typ[i] == 'I') {
			for (int64_t j = 0; j < cnt[i]; j++, index++)
				if (i != 0 && i != cnt.size() - 1) o << query.at(index);
		} else if (typ[i] != 'D') {
			index += (size_t) cnt[i];
		}
	}
	o << "\n";
	return o.str();
}

// Under -v: SubjectRead::writeAlignmentLongFormat per subject (SubjectRead.cpp:195-206) — ">name\tlength",
// a line per record, an empty line — and ">-1" with an empty line for a subject that is not good
// (SubjectReadAligner.cpp:920-952, where the uniqueId under -v is unset).
void SubjectReadAligner::takeRecords(const ReadBatch &subjects, const std::vector<uint8_t> &good) {
	Device &dev = query_->device();
	int64_t nr = 0, no = 0;
	dev.check(hero_take_records(dev.ctx(), nullptr, 0, &nr, nullptr, 0, &no), "hero_take_records");
	std::vector<hero_record> recs((size_t) nr);
	std::vector<uint32_t> ops((size_t) no);
	dev.check(hero_take_records(dev.ctx(), recs.data(), nr, &nr, ops.data(), no, &no), "hero_take_records");
	const ReadBatch &Q = query_->reads();
	size_t x = 0;
	for (int64_t j = 0; j < subjects.size(); j++) {
		if (!good[(size_t) j]) {
			aln_ += ">-1\n\n";
			while (x < recs.size() && recs[x].subject == j) x++;
			continue;
		}
		aln_ += ">" + (subjects.name.empty() ? std::string() : subjects.name[(size_t) j]) + "\t" +
				std::to_string(subjects.length[(size_t) j]) + "\n";
		for (; x < recs.size() && recs[x].subject == j; x++) {
			const hero_record &r = recs[x];
			const std::string q = Q.read(r.query);
			const std::string name = Q.name.empty() ? std::to_string((long long) r.query) : Q.name[(size_t) r.query];
			aln_ += longFormatRecord(name, r.strand == 1 ? rc_of(q) : q, r.strand, r.position, ops.data() + r.op_off,
					r.op_n, subjects.length[(size_t) j]);
		}
		aln_ += "\n";
	}
}

bool SubjectReadAligner::start(SubjectDataset &subjects) {
	Device &dev = query_->device();
	while (subjects.loadNextChunkParallel()) {
		const ReadBatch &c = subjects.chunk;
		dev.check(hero_align_subjects(dev.ctx(), c.size(), c.seq.data(), c.offset.data(), c.length.data(),
						  c.unique_id.data()),
				"hero_align_subjects");
		if (dev.options().call_consensus) continue;
		if (dev.options().long_format) takeRecords(c, subjects.good);
		else take_alignments(dev, aln_);
	}
	return true;
}

bool SubjectReadAligner::start() {
	Device &dev = query_->device();
	const Config &cfg = dev.config();
	SubjectDataset subjects(cfg);
	aln_.clear();
	const bool write = !dev.options().call_consensus;
	while (subjects.loadNextChunkParallel()) {
		const ReadBatch &c = subjects.chunk;
		dev.check(hero_align_subjects(dev.ctx(), c.size(), c.seq.data(), c.offset.data(), c.length.data(),
						  c.unique_id.data()),
				"hero_align_subjects");
		if (!write) continue;
		if (dev.options().long_format) takeRecords(c, subjects.good);
		else take_alignments(dev, aln_);
		// SubjectReadAligner::writeToFile[2] (SubjectReadAligner.cpp:854-960): appended per chunk
		std::ofstream(cfg.outputfilename + ".aln", std::ios_base::app | std::ios_base::binary) << aln_;
		aln_.clear();
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
