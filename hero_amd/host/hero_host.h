// C++ host classes with the shapes of the reference's hot-path seams, over the C ABI
// (include/hero_gpu.h). A caller of guo-xuan/Hero's QueryDataset / SubjectDataset /
// HashTableMethod / BandAligner / SubjectReadAligner finds the same names, argument meanings and
// error behaviour here; the work runs in libhero_gpu on the device.
//
// Reference interfaces mirrored (guo-xuan/Hero, src/):
//   QueryDataset        QueryDataset.h:14-52
//   SubjectDataset      SubjectDataset.h:19-39
//   HashTableLongKey    HashTableMethod.h:13-40, HashTableLongKey.h
//   BandAligner         BandAligner.h:20-92
//   SubjectReadAligner  SubjectReadAligner.h (start(), SubjectReadAligner.cpp:493-645)
// File parsing is limited to two-line FASTA with integer names (uniqueId = atoi(name)).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/hero_gpu.h"

namespace hero_host {

// Config: the run options (Config.cpp:17-56 defaults) and the device context they configure.
class Config {
public:
	hero_options opt;
	Config();
	// -k -m -i -l -a (g|a) -c flags of the reference's command line, as the options they set
	void setKmer(int k) { opt.k = k; }
	void setMaxMismatchRatePercent(double m) { opt.mismatch_rate = m / 100.0; }
	void setMaxIndelRatePercent(double i) { opt.indel_rate = i / 100.0; }
	void setMinimumOverlap(int64_t l) { opt.min_overlap = l; }
	void setAffine(bool affine) { opt.aligner = affine ? 0 : 1; }
	void setCallConsensus(bool c) { opt.call_consensus = c ? 1 : 0; }
	void setMapping(bool m) { opt.mapping = m ? 1 : 0; } // --mapping (Config::bMappingVersion)
};

// The device context; every dataset / aligner object borrows one.
class Device {
public:
	explicit Device(const Config &cfg);
	~Device();
	Device(const Device &) = delete;
	Device &operator=(const Device &) = delete;
	hero_ctx *ctx() const { return ctx_; }
	const hero_options &options() const { return opt_; }
	// Throws std::runtime_error with hero_last_error() on a failed call.
	void check(int rc, const char *what) const;

private:
	hero_ctx *ctx_ = nullptr;
	hero_options opt_;
};

// Reads as the datasets hold them: ASCII sequences back to back, per-read offset/length, uniqueId.
struct ReadBatch {
	std::string seq;
	std::vector<int64_t> offset;
	std::vector<int32_t> length;
	std::vector<int64_t> unique_id;
	int64_t size() const { return (int64_t) length.size(); }
	void add(const std::string &s, int64_t uid);
	std::string read(int64_t i) const { return seq.substr((size_t) offset[(size_t) i], (size_t) length[(size_t) i]); }
	// two-line FASTA, names are integers (the self-hit skip compares atoi(name))
	static bool loadFasta(const std::string &path, ReadBatch &out);
};

class QueryDataset {
public:
	uint64_t longestReadLength = 0;
	QueryDataset(Device &dev, const ReadBatch &reads);
	static bool qualityFilter(const std::string &sequence); // QueryDataset.cpp:60-75
	uint64_t getNumberOfReads() const { return (uint64_t) reads_.size(); }
	// The read's current sequence, or nullptr when out of range (QueryDataset.cpp:49-58).
	const std::string *getReadFromID(int64_t id);
	// QueryDataset::callConsensus (QueryDataset.cpp:393-412): the corrected reads replace the reads.
	void callConsensus();
	// Corrected FASTA as QueryDataset::writeQueryOut writes it without --merge (QueryDataset.cpp:414-487).
	void writeQueryOut(const std::string &filename);
	std::string queryOutText() const;
	// --mapping output (QueryDataset::getCoverageAndWriteResults, QueryDataset.cpp:560-600): the
	// bases by coverage, appended to the file as "COV\tFRE" and "coverage\tbases" lines.
	void getCoverageAndWriteResults(const std::string &filename);
	std::string coverageText() const;
	Device &device() const { return dev_; }
	const ReadBatch &reads() const { return reads_; }

private:
	Device &dev_;
	ReadBatch reads_;
	std::vector<std::string> cache_;
};

// HashTableMethod / HashTableLongKey over the device index of the query reads.
class HashTableLongKey {
public:
	QueryDataset *dataSet = nullptr;
	int64_t hashKeyLength;
	explicit HashTableLongKey(QueryDataset *q);
	bool createHashTables() { return true; } // sizing happens in insertQueryDataset
	bool insertQueryDataset(QueryDataset *q); // HashTableLongKey.cpp:250-306
	// Entries ((pos << 40) | innerId, insertion order) of the k-mer at _sequence[_start, _start + _length);
	// empty when the k-mer is absent (HashTableLongKey.cpp:79-99). The reference returns a borrowed
	// pointer; this returns a copy.
	std::vector<int64_t> getListOfReads(const std::string &_sequence, int64_t _start, int64_t _length);
	// hashFunction (HashTableLongKey.cpp:188-206) for table size P.
	uint64_t hashFunction(const std::string &_subString, uint64_t P);

private:
	Device &dev_;
};

// BandAligner: setBand -> setSequences -> setAlignLeft / setAlignRight -> calculateAlignment ->
// getCigarType / getCigarCount. One call is one device batch of one pair; BandAlignerBatch runs
// many pairs per device call.
class BandAligner {
public:
	std::vector<int> vCigarCount;
	std::vector<char> vCigarType;
	int64_t m_band = 0;
	explicit BandAligner(Device &dev) : dev_(dev) {}
	void setBand(int64_t band) { m_band = band; }
	void setAlignLeft() { flag_ = 1; }
	void setAlignRight() { flag_ = 2; }
	// end exclusive; exits the process on a negative start, as the reference (BandAligner.cpp:546-549,577-580)
	void setSequences(const std::string &S, int64_t s0, int64_t s1, const std::string &Q, int64_t q0, int64_t q1);
	void calculateAlignment();
	std::vector<int> *getCigarCount() { return &vCigarCount; }
	std::vector<char> *getCigarType() { return &vCigarType; }

private:
	Device &dev_;
	std::string s_, q_;
	int flag_ = 0;
};

struct BandPairResult {
	int64_t band;
	std::vector<int> counts;
	std::vector<char> types;
};
class BandAlignerBatch {
public:
	explicit BandAlignerBatch(Device &dev) : dev_(dev) {}
	void add(const std::string &S, int64_t s0, int64_t s1, const std::string &Q, int64_t q0, int64_t q1, int64_t band,
			int flag);
	std::vector<BandPairResult> run();

private:
	Device &dev_;
	std::string seq_;
	std::vector<hero_band_pair> pairs_;
};

// Chunks of subject reads (SubjectDataset::loadNextChunkParallel, SubjectDataset.cpp:45-179).
class SubjectDataset {
public:
	ReadBatch chunk;
	uint64_t iCountSubjectRead = 0;
	explicit SubjectDataset(const ReadBatch &all, int64_t chunk_size = 400) : all_(all), chunk_size_(chunk_size) {}
	bool loadNextChunkParallel(uint16_t threads = 1); // false at the end
	bool clearSubjectRead() {
		chunk = ReadBatch();
		return true;
	}

private:
	const ReadBatch &all_;
	int64_t chunk_size_;
	int64_t next_ = 0;
};

// SubjectReadAligner::start: every subject chunk against the query index; votes accumulate on
// the device (-c), or the .aln text is returned per chunk (without -c).
class SubjectReadAligner {
public:
	SubjectReadAligner(QueryDataset *q, HashTableLongKey *table) : query_(q), table_(table) {}
	bool start(SubjectDataset &subjects);
	// self-correction (-q f -s f): the query set is the subject stream, no second upload
	bool startSelf();
	const std::string &alignments() const { return aln_; }

private:
	QueryDataset *query_;
	HashTableLongKey *table_;
	std::string aln_;
};

} // namespace hero_host
