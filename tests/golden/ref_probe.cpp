// Known-answer probe over the reference's own classes.
//
// Built by oracle/Makefile (target `ref`) against objects compiled from the reference sources
// where they lie; this file is the only code here and it only *calls* the reference API:
//   BandAligner      (BandAligner.h:20-92)  setBand / setSequences / setAlignLeft / setAlignRight /
//                                           calculateAlignment / getCigarType / getCigarCount, m_band
//   HashTableLongKey (HashTableLongKey.h)   hashFunction(string), getPrimeLargerThanNumber
//   AlignmentRecord  (AlignmentRecord.h)    addCigar / isGoodAlignment / calculateScore fields
//   QueryRead        (QueryRead.h)          setSequence / initilizeCoverageMatrix / callConsensus
//
// Protocol: one command per stdin line, one answer line per command on stdout.
//   T <0|1>                       aligner type (0 affine, 1 global); rebuilds the BandAligner
//   B <band>                      setBand(band)
//   D <S|-> <Q|-> <flag>          setSequences(S,0,|S|,Q,0,|Q|); flag 1=setAlignLeft 2=setAlignRight;
//                                 calculateAlignment(); prints "<m_band> <count><type>..."
//   H <P> <kmer>                  prints hashFunction(kmer) with hashTableSize=P
//   P <n>                         prints getPrimeLargerThanNumber(n)
//   R <mis> <indel> <cnt><type>.. prints "<nMis> <nIndel> <len> <good_caps> <good_rates>" after
//                                 addCigar of every op (rates: -m/-i as fractions)
//   C <thr> <seq> <v0> <v1> ...   sets the 4*len tally, calls callConsensus(), prints the sequence
// The Config defaults that matter are set explicitly: iLongestRead=1100 (DP scratch size).
#include "Config.h"
#include "BandAligner.h"
#include "AlignmentRecord.h"
#include "HashTableLongKey.h"
#include "QueryRead.h"

#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

static std::string cigar_string(BandAligner* a) {
	std::ostringstream os;
	std::vector<char>* t = a->getCigarType();
	std::vector<int>* c = a->getCigarCount();
	for (size_t i = 0; i < t->size(); i++) os << c->at(i) << t->at(i);
	return os.str();
}

int main() {
	std::ios::sync_with_stdio(false);
	Config::iLongestRead = 1100;
	Config::iAlignerType = BandAligner::IGLOBAL;
	BandAligner* aligner = new BandAligner();
	HashTableLongKey* table = new HashTableLongKey(NULL);
	std::string line;
	while (std::getline(std::cin, line)) {
		if (line.empty()) continue;
		std::istringstream in(line);
		std::string cmd;
		in >> cmd;
		if (cmd == "T") {
			int t;
			in >> t;
			delete aligner;
			Config::iAlignerType = (UINT16) t;
			aligner = new BandAligner();
			std::cout << "ok\n";
		} else if (cmd == "B") {
			long long b;
			in >> b;
			aligner->setBand(b);
			std::cout << "ok\n";
		} else if (cmd == "D") {
			std::string s, q;
			int flag;
			in >> s >> q >> flag;
			if (s == "-") s = "";
			if (q == "-") q = "";
			aligner->setSequences(s, 0, (INT64) s.size(), q, 0, (INT64) q.size());
			if (flag == 1) aligner->setAlignLeft();
			if (flag == 2) aligner->setAlignRight();
			aligner->calculateAlignment();
			std::cout << aligner->m_band << " " << cigar_string(aligner) << "\n";
		} else if (cmd == "H") {
			unsigned long long p;
			std::string kmer;
			in >> p >> kmer;
			table->hashTableSize = p;
			std::cout << table->hashFunction(kmer) << "\n";
		} else if (cmd == "P") {
			unsigned long long n;
			in >> n;
			std::cout << table->getPrimeLargerThanNumber(n) << "\n";
		} else if (cmd == "R") {
			double mis_rate, indel_rate;
			long long cap_mis, cap_indel;
			in >> mis_rate >> indel_rate >> cap_mis >> cap_indel;
			Config::dMaxMismatchRate = mis_rate;
			Config::dMaxIndelRate = indel_rate;
			AlignmentRecord rec;
			std::string op;
			while (in >> op) {
				char type = op[op.size() - 1];
				long long cnt = atoll(op.substr(0, op.size() - 1).c_str());
				rec.addCigar(cnt, type);
			}
			INT64 cm = cap_mis, ci = cap_indel;
			bool good_caps = rec.isGoodAlignment(cm, ci);
			bool good_rates = rec.isGoodAlignment();
			char buf[256];
			snprintf(buf, sizeof buf, "%.17g %.17g %.17g %d %d %.17g", rec.dNumMismatch, rec.dNumIndel,
					rec.dAlignmentLength, (int) good_caps, (int) good_rates, rec.dIndelRate + rec.dMismatchRate);
			std::cout << buf << " ";
			std::ostringstream os;
			for (size_t i = 0; i < rec.vCigarType.size(); i++) os << rec.vCigarCount[i] << rec.vCigarType[i];
			std::cout << os.str() << "\n";
		} else if (cmd == "C") {
			double thr;
			std::string seq;
			in >> thr >> seq;
			Config::dMinFrequencyCallConsensus = thr;
			Config::bCallConsensus = true;
			QueryRead r;
			r.setSequence(seq);
			r.initilizeCoverageMatrix();
			for (size_t i = 0; i < seq.size() * 4; i++) {
				unsigned long v;
				in >> v;
				r.aCoverageMatrix[i] = (UINT32) v;
			}
			r.callConsensus();
			std::cout << r.sReadSequence << "\n";
		} else {
			std::cout << "error unknown command\n";
		}
		std::cout.flush();
	}
	delete aligner;
	return 0;
}
