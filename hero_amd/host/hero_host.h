// C++ host classes with the shapes of the reference's hot-path seams, over the C ABI
// (include/hero_gpu.h). A caller of guo-xuan/Hero's QueryDataset / SubjectDataset /
// HashTableMethod / BandAligner / SubjectReadAligner finds the same names, argument meanings and
// error behaviour here; the work runs in libhero_gpu on the device.
//
// Reference interfaces mirrored (guo-xuan/Hero, src/):
//   Config              Config.h, Config.cpp:11-56 (the run options and file lists)
//   QueryDataset        QueryDataset.h:14-52 (FASTA / FASTQ loading, consensus, output files)
//   SubjectDataset      SubjectDataset.h:19-39 (streamed subject chunks)
//   HashTableMethod     HashTableMethod.h:13-40, HashTableLongKey.h
//   BandAligner         BandAligner.h:20-92
//   SubjectReadAligner  SubjectReadAligner.h:31-41 (start(), SubjectReadAligner.cpp:493-645)
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hero_gpu.h"

namespace hero_host {

// Config: the run options (Config.cpp:17-56 defaults) and the files of a run.
class Config {
public:
	hero_options opt;
	// -q: one file (sQueryFileName) or a comma-separated list (vQueryFiles); -s: subject files; -o
	std::string sQueryFileName;
	std::vector<std::string> vQueryFiles, vSubjectFiles;
	std::string outputfilename = "query";
	uint16_t numberOfThreads = 1;     // -t
	uint64_t streamChunkSize = 400;   // -z: subject reads per streamed chunk
	uint64_t queryChunkSize = 800;    // -y: query reads per loading chunk
	bool bSpeicialVersion = false;    // -v: names kept, long-format .aln
	bool bMergedReads = false;        // --singleReads: corrected reads written one by one, not in pairs
	bool bCoverageStatistic = false;  // --covinfo: the .info coverage file
	bool bMerging = false;            // --merge: paired reads merged after consensus
	uint32_t iCoverageDepthFilter = 1; // --covfilter (read by QueryRead::isGoodRead, which returns first)
	Config();
	// -k -m -i -l -a (g|a) -c -f flags of the reference's command line, as the options they set
	void setKmer(int k) { opt.k = k; }
	void setMaxMismatchRatePercent(double m) { opt.mismatch_rate = m / 100.0; }
	void setMaxIndelRatePercent(double i) { opt.indel_rate = i / 100.0; }
	void setMinimumOverlap(int64_t l) { opt.min_overlap = l; }
	void setAffine(bool affine) { opt.aligner = affine ? 0 : 1; }
	void setCallConsensus(bool c) { opt.call_consensus = c ? 1 : 0; }
	void setConsensusFrequency(double f) { opt.consensus_freq = f; }
	void setHashTableFactor(int f) { opt.hash_factor = f; }
	void setMapping(bool m) { opt.mapping = m ? 1 : 0; } // --mapping (Config::bMappingVersion)
	void setSpecialVersion(bool v) { bSpeicialVersion = v; }
	void setMerging(bool m) { bMerging = m; }
	// Sets the options every run derives from the flags: -v and --merge in the device options, and
	// record mode (.aln) without -c.
	hero_options deviceOptions() const;
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
	const Config &config() const { return cfg_; }
	// Throws std::runtime_error with hero_last_error() on a failed call.
	void check(int rc, const char *what) const;

private:
	hero_ctx *ctx_ = nullptr;
	Config cfg_;
	hero_options opt_;
};

// Reads as the datasets hold them: ASCII sequences back to back, per-read offset/length, uniqueId
// and (with -v) the read's name.
struct ReadBatch {
	std::string seq;
	std::vector<int64_t> offset;
	std::vector<int32_t> length;
	std::vector<int64_t> unique_id;
	std::vector<std::string> name;
	int64_t size() const { return (int64_t) length.size(); }
	void add(const std::string &s, int64_t uid);
	std::string read(int64_t i) const { return seq.substr((size_t) offset[(size_t) i], (size_t) length[(size_t) i]); }
	// the reads of FASTA / FASTQ files as QueryDataset::loadDatasetParallel reads them (uniqueId =
	// atoi(name)); false when a file cannot be opened
	static bool loadFasta(const std::string &path, ReadBatch &out);
};

// FASTA / FASTQ records of a list of files, read as the reference's loaders read them
// (QueryDataset.cpp:103-391, SubjectDataset.cpp:45-179): the format comes from the first file's
// first line ('>' FASTA, '@' FASTQ); a FASTA record is its header line and everything up to the
// next '>' with newlines removed; a FASTQ record is four lines; the name is the header without its
// leading '>' / '@'; sequences are upper-cased; a last record whose sequence (FASTA) or fourth line
// (FASTQ) is empty is dropped.
class ReadFiles {
public:
	explicit ReadFiles(const std::vector<std::string> &files);
	~ReadFiles();
	bool ok() const { return ok_; }       // every file opened
	// the next record; false at the end of the last file
	bool next(std::string &name, std::string &sequence);
	// the subject streamer's step (SubjectDataset.cpp:79-121): at a file's end open the next file
	// (false after the last), then read one record; `text` is its last read, which the streamer
	// tests at a chunk's end
	bool nextRaw(std::string &name, std::string &sequence, std::string &text);

private:
	struct Impl;
	std::unique_ptr<Impl> p_;
	bool ok_ = true;
};

class QueryDataset {
public:
	uint64_t longestReadLength = 0;
	// QueryDataset(const string &) / QueryDataset(const vector<string> &) + buildDataset()
	// (QueryDataset.cpp:10-33,90-102): the file(s) are read by buildDataset / loadDatasetParallel.
	QueryDataset(Device &dev, const std::string &queryFilename);
	QueryDataset(Device &dev, const std::vector<std::string> &queryFilenames);
	// reads already in memory (loaded at once)
	QueryDataset(Device &dev, const ReadBatch &reads);
	bool buildDataset();
	bool loadDatasetParallel(const std::string &filename, uint16_t numberOfThreads);
	bool loadDatasetParallel(const std::vector<std::string> &filenames, uint16_t numberOfThreads);
	static bool qualityFilter(const std::string &sequence); // QueryDataset.cpp:60-75
	uint64_t getNumberOfReads() const { return (uint64_t) reads_.size(); }
	// The read's current sequence, or nullptr when out of range (QueryDataset.cpp:49-58).
	const std::string *getReadFromID(int64_t id);
	// QueryDataset::callConsensus (QueryDataset.cpp:393-412): the corrected reads replace the reads
	// (and, with --covinfo, each good read's coverage statistics, QueryRead.cpp:192-212)
	void callConsensus();
	// QueryDataset::mergePairedReads (QueryDataset.cpp:539-558, QueryRead.cpp:268-317): reads 2i and
	// 2i+1 are mates; the mate's reverse complement extends the left read where they overlap and the
	// extension tallies agree
	void mergePairedReads();
	// QueryDataset::writeQueryOut (QueryDataset.cpp:414-537): <name>.fasta (reads in pairs, or one by
	// one with --singleReads), or <name>_merged.fasta / <name>_unmerged.fasta with --merge; and
	// <name>.info with --covinfo. Files are appended to, as the reference does.
	void writeQueryOut(const std::string &filename);
	std::string queryOutText() const; // the --singleReads FASTA text
	// --mapping output (QueryDataset::getCoverageAndWriteResults, QueryDataset.cpp:560-600): the
	// bases by coverage, appended to the file as "COV\tFRE" and "coverage\tbases" lines.
	void getCoverageAndWriteResults(const std::string &filename);
	std::string coverageText() const;
	Device &device() const { return dev_; }
	const ReadBatch &reads() const { return reads_; }
	// per read: QueryRead::bIsGoodRead (len > minOverlap && qualityFilter) and bMerged
	bool isGood(int64_t i) const { return good_[(size_t) i] != 0; }
	bool isMerged(int64_t i) const { return merged_[(size_t) i] != 0; }

private:
	struct Stat {
		uint32_t iMin = 0, iMed = 0, iMax = 0;
		double dActualMismatchRate = 0, dGcContent = 0;
	};
	void upload();
	std::string infoLine(int64_t i) const;
	void writeRecord(std::string &out, int64_t i) const;
	Device &dev_;
	std::vector<std::string> files_;
	ReadBatch reads_;
	std::vector<uint32_t> raw_len_;   // QueryRead::iRawReadLength
	std::vector<uint8_t> good_, merged_;
	std::vector<Stat> stat_;
	std::vector<std::string> cache_;
	bool loaded_ = false;
};

// HashTableMethod (HashTableMethod.h:13-40): the query index's interface. The public data mirrors
// the reference's (callers read iMask / iBitOfReadId / dataSet); the table itself lives on the
// device.
class HashTableMethod {
public:
	QueryDataset *dataSet = nullptr;
	uint64_t hashTableSize = 0;
	int64_t iBitOfReadId = 40;
	int64_t iMask = (1ll << 40) - 1;
	int64_t hashKeyLength = 0;
	int64_t numberOfHashCollision = 0;
	uint64_t maxSingleHashCollision = 0;
	virtual ~HashTableMethod() {}
	virtual bool createHashTables() = 0;
	virtual std::vector<int64_t> *getListOfReads(const std::string &sequence, int64_t start, int64_t length) = 0;
	virtual uint64_t getPrimeLargerThanNumber(uint64_t number) = 0;
	virtual uint64_t hashFunction(const std::string &subString) = 0;
	virtual uint64_t hashFunction(const std::string &sequence, int64_t start, int64_t length) = 0;
	virtual bool insertQueryDataset(QueryDataset *qDataset) = 0;
	virtual bool insertQueryRead(int64_t readInnerId, const std::string &subString, int64_t position) = 0;
	virtual bool isEmptyAt(int64_t hashTableIndex) = 0;
	virtual void setHashTableSizeAndInitialize(int64_t size) = 0;
};

// HashTableLongKey (HashTableLongKey.cpp) over the device index of the query reads.
class HashTableLongKey : public HashTableMethod {
public:
	explicit HashTableLongKey(QueryDataset *q);
	// size = the smallest listed prime above factor x reads (factor x bases / k under --mapping),
	// HashTableLongKey.cpp:41-68
	bool createHashTables() override;
	// Entries ((pos << iBitOfReadId) | innerId, insertion order) of the k-mer at
	// sequence[start, start + length), as a pointer the table owns and keeps valid for its lifetime;
	// nullptr when the k-mer is absent (HashTableLongKey.cpp:79-99).
	std::vector<int64_t> *getListOfReads(const std::string &sequence, int64_t start, int64_t length) override;
	uint64_t getPrimeLargerThanNumber(uint64_t number) override; // HashTableLongKey.cpp:101-166
	// hashFunction (HashTableLongKey.cpp:168-206) for the table's hashTableSize
	uint64_t hashFunction(const std::string &subString) override;
	uint64_t hashFunction(const std::string &sequence, int64_t start, int64_t length) override;
	uint64_t hashFunction(const std::string &subString, uint64_t P); // for a given table size
	bool insertQueryDataset(QueryDataset *qDataset) override; // HashTableLongKey.cpp:250-306, on the device
	// The device index is built from every good read's keys at once (insertQueryDataset); single
	// insertions are not taken: returns false.
	bool insertQueryRead(int64_t readInnerId, const std::string &subString, int64_t position) override;
	// Whether slot i of the reference's table (linear probing from each key's hash, in insertion
	// order) is empty; maxSingleHashCollision and numberOfHashCollision are the reference's counts
	// of that same insertion, computed with it on the first call.
	bool isEmptyAt(int64_t hashTableIndex) override;
	void setHashTableSizeAndInitialize(int64_t size) override;

private:
	void placeKeys();
	Device &dev_;
	std::unordered_map<std::string, std::unique_ptr<std::vector<int64_t>>> lists_;
	std::vector<uint64_t> occupied_; // bitmap of the reference's occupied slots (placeKeys)
	bool placed_ = false;
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

// Chunks of subject reads (SubjectDataset::loadNextChunkParallel, SubjectDataset.cpp:45-179): from the
// configured subject files (streamChunkSize reads per chunk, files read one after the other), or
// from reads in memory. Each chunk's reads carry uniqueId = atoi(name) (-1 under -v, where the name
// is kept) and the subject predicate good = len < minOverlap || qualityFilter.
class SubjectDataset {
public:
	ReadBatch chunk;
	std::vector<uint8_t> good; // SubjectRead::bIsGoodRead of the chunk's reads
	uint64_t iCountSubjectRead = 0;
	uint64_t iLongestReadLength = 0;
	explicit SubjectDataset(const Config &cfg);                       // Config::vSubjectFiles
	SubjectDataset(const ReadBatch &all, int64_t chunk_size = 400); // reads in memory
	~SubjectDataset();
	bool loadNextChunkParallel(uint16_t threads = 1); // false at the end
	bool clearSubjectRead();

private:
	const ReadBatch *all_ = nullptr;
	std::unique_ptr<ReadFiles> files_;
	int64_t chunk_size_;
	int64_t next_ = 0;
	int64_t min_overlap_ = 40;
	bool keep_names_ = false;
};

// SubjectReadAligner::start: every subject chunk against the query index; votes accumulate on
// the device (-c), or the alignment records are written to <output>.aln (the reference's format,
// or the long format under -v).
class SubjectReadAligner {
public:
	// SubjectReadAligner(HashTableMethod *) + start() (SubjectReadAligner.h:34-42): the subjects are
	// the Config's subject files, written to Config::outputfilename + ".aln" without -c
	explicit SubjectReadAligner(HashTableMethod *table);
	SubjectReadAligner(QueryDataset *q, HashTableMethod *table) : query_(q), table_(table) {}
	bool start();
	bool start(SubjectDataset &subjects); // the text of .aln kept in memory (alignments())
	// self-correction (-q f -s f): the query set is the subject stream, no second upload
	bool startSelf();
	const std::string &alignments() const { return aln_; }

private:
	void takeRecords(const ReadBatch &subjects, const std::vector<uint8_t> &good);
	QueryDataset *query_;
	HashTableMethod *table_;
	std::string aln_;
};

// AlignmentRecord's long-format fields (AlignmentRecord.cpp:151-381, SubjectRead.cpp:102-121,
// 195-206): one record line of SubjectRead::getAlignmentLongFormat for a record of ops
// ((count << 8) | type) at position pos, query read `query` in the record's orientation.
std::string longFormatRecord(const std::string &queryName, const std::string &query, int strand, int64_t pos,
		const uint32_t *ops, int nops, int64_t subjectLength);

} // namespace hero_host
