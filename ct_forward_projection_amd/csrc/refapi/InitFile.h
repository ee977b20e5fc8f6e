// InitFile.h -- INI reader with the interface of the reference's CInitFile
// (Shared/Include/InitFile.h:40-104), written from that declaration and the
// shipped Config/*.ini files only (the reference implementation is a closed
// binary and was not used).
//
// Accepted grammar: "[Section]" lines; "key=value" lines; a value ends at the
// first ';' (the rest of the line is a comment); surrounding spaces/tabs are
// trimmed; '#' or ';' at line start comments the line; UTF-8 text; CRLF or LF.
// Section and key names match exactly, else case-insensitively.
#pragma once
#include <cstddef>
#include <map>
#include <string>
#include <vector>

typedef int BOOL;

class CInitFile {
  public:
    CInitFile();
    ~CInitFile();

    // Loads and parses `szFileName`; a missing file leaves the reader empty
    // (every lookup then returns its default).  IsLoaded() reports which.
    void GetFileName(const char *szFileName = "Reconstruct.txt");
    bool IsLoaded() const { return m_loaded; }
    std::string GetParaFilePath() const { return m_szModName; }

    // Writes the current sections back (to szFileName, or the loaded file).
    void SaveFile(const char *szFileName = 0);
    void creatFile(const char *szFileName);

    BOOL SetEntryValue(const char *szSection, const char *szEntry, std::string szValue);
    BOOL SetEntryValue(const char *szSection, const char *szEntry, char iValue);
    BOOL SetEntryValue(const char *szSection, const char *szEntry, short iValue);
    BOOL SetEntryValue(const char *szSection, const char *szEntry, int iValue);
    BOOL SetEntryValue(const char *szSection, const char *szEntry, unsigned long iValue);
    BOOL SetEntryValue(const char *szSection, const char *szEntry, double dbValue);
    BOOL SetEntryValue(const char *szSection, const char *szEntry, long long i64value);
    BOOL SetEntryValue(const char *szSection, const char *szEntry, float iValue);

    // Each returns TRUE and parses the entry when present and well formed;
    // otherwise stores the default and returns FALSE.
    BOOL GetEntryValue(const char *szSection, const char *szEntry, const char *szDefault,
                       std::string &szValue);
    BOOL GetEntryValue(const char *szSection, const char *szEntry, int iDefault, char &iValue);
    BOOL GetEntryValue(const char *szSection, const char *szEntry, int iDefault, short &iValue);
    BOOL GetEntryValue(const char *szSection, const char *szEntry, int iDefault, int &iValue);
    BOOL GetEntryValue(const char *szSection, const char *szEntry, long iDefault, long &iValue);
    BOOL GetEntryValue(const char *szSection, const char *szEntry, double dbDefault, double &dbValue);
    BOOL GetEntryValue(const char *szSection, const char *szEntry, long long i64Default,
                       long long &i64Value);
    BOOL GetEntryValue(const char *szSection, const char *szEntry, size_t iDefault, size_t &iValue);
    BOOL GetEntryValue(const char *szSection, const char *szEntry, float iDefault, float &iValue);

  private:
    struct Entry {
        std::string key, value, comment;
    };
    struct Section {
        std::string name;
        std::vector<Entry> entries;
    };
    const Entry *find(const char *section, const char *key) const;
    Entry *findOrAdd(const char *section, const char *key);
    void parse(const std::string &text);

    std::string m_szModName;
    std::vector<Section> m_sections;
    bool m_loaded = false;
};
